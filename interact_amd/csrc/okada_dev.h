// Device-side Okada (1992) rectangular dislocation: displacement GRADIENTS
// only, for unit strike slip and unit dip slip at once.
//
// The operator needs stress, so the displacement terms of parts A and B
// (and with them every atan / log of the solution: TT, ALX, ALE, I1..I4)
// drop out; part C's displacement terms stay because they enter the z
// derivative of the z*u_C product.  Both slip directions share each
// corner's geometry.  The 1/(2 pi) factor is applied once at the end.
//
// Behavioural guards of the reference's DC3D are kept
// (src/green/dc3d.F:646-813): 1e-8 snapping of xi, eta, q and cos(dip);
// edge-singular observation points return iret = 1 (entry 0); the
// negative-extension flags KXI/KET select the R-xi limit forms; full space
// keeps the real-source term only.
#pragma once
#include <hip/hip_runtime.h>

namespace rsfg {

constexpr double kOkEps = 1.0e-8;
constexpr double kInv2Pi = 0.15915494309189533577;  // 1 / (2 pi)

struct OkMedium {
  double a1, a2, a3, a4, a5;  // (1-alpha)/2, alpha/2, (1-alpha)/alpha, 1-alpha, alpha
};

__device__ __forceinline__ double ok_snap(double v) {
  return fabs(v) < kOkEps ? 0.0 : v;
}

// gradient 9-vectors are ordered ux,x uy,x uz,x ux,y uy,y uz,y ux,z uy,z uz,z
struct OkGrad {
  double s[9];  // unit strike slip
  double d[9];  // unit dip slip
};

// One corner (xi, et) of the rectangle for the real (IMAGE = false) or
// image (IMAGE = true) source, accumulated into g with the corner sign.
template <bool IMAGE>
__device__ __forceinline__ void ok_corner(double xi, double et, double q,
                                          bool kxi, bool ket, double sd,
                                          double cd, double z,
                                          const OkMedium& m, double sgn,
                                          OkGrad& g) {
  const double xi2 = xi * xi, et2 = et * et, q2 = q * q;
  const double r2 = xi2 + et2 + q2;
  const double r = sqrt(r2);
  const double ir = 1.0 / r;
  const double ir2 = ir * ir;
  const double ir3 = ir * ir2;
  const double yb = et * cd + q * sd;  // y-bar
  const double db = et * sd - q * cd;  // d-bar
  double x11 = 0.0, x32 = 0.0, y11 = 0.0, y32 = 0.0;
  if (!kxi) {
    const double rx = r + xi;
    x11 = 1.0 / (r * rx);
    x32 = (r + rx) * x11 * x11 * ir;
  }
  if (!ket) {
    const double re = r + et;
    y11 = 1.0 / (r * re);
    y32 = (r + re) * y11 * y11 * ir;
  }
  const double qr3 = q * ir3;
  const double ey = sd * ir - yb * qr3;
  const double ez = cd * ir + db * qr3;
  const double fy = db * ir3 + xi2 * y32 * sd;
  const double fz = yb * ir3 + xi2 * y32 * cd;
  const double gy = 2.0 * x11 * sd - yb * q * x32;
  const double gz = 2.0 * x11 * cd + db * q * x32;
  const double xy = xi * y11, qy = q * y11;

  // part A (infinite-medium) gradients
  double as[9], ad[9];
  as[0] = -m.a1 * qy - m.a2 * q * xi2 * y32;
  as[1] = -m.a2 * xi * qr3;
  as[2] = m.a1 * xy + m.a2 * xi * q2 * y32;
  as[3] = m.a1 * xy * sd + m.a2 * xi * fy + 0.5 * db * x11;
  as[4] = m.a2 * ey;
  as[5] = m.a1 * (cd * ir + qy * sd) - m.a2 * q * fy;
  as[6] = m.a1 * xy * cd + m.a2 * xi * fz + 0.5 * yb * x11;
  as[7] = m.a2 * ez;
  as[8] = -m.a1 * (sd * ir - qy * cd) - m.a2 * q * fz;
  ad[0] = -m.a2 * xi * qr3;
  ad[1] = -0.5 * qy - m.a2 * et * qr3;
  ad[2] = m.a1 * ir + m.a2 * q2 * ir3;
  ad[3] = m.a2 * ey;
  ad[4] = m.a1 * x11 * db + 0.5 * xy * sd + m.a2 * et * gy;
  ad[5] = m.a1 * x11 * yb - m.a2 * q * gy;
  ad[6] = m.a2 * ez;
  ad[7] = m.a1 * x11 * yb + 0.5 * xy * cd + m.a2 * et * gz;
  ad[8] = -m.a1 * x11 * db - m.a2 * q * gz;

  if (!IMAGE) {
    // real source: reflected frame; x and y derivative groups flip sign
    // of the first component and rotate the other two, z group is the
    // negative of that
#pragma unroll
    for (int grp = 0; grp < 3; ++grp) {
      const int i = 3 * grp;
      const double zs = (grp == 2) ? -1.0 : 1.0;
      g.s[i] += sgn * zs * (-as[i]);
      g.s[i + 1] += sgn * zs * (-as[i + 1] * cd + as[i + 2] * sd);
      g.s[i + 2] += sgn * zs * (-as[i + 1] * sd - as[i + 2] * cd);
      g.d[i] += sgn * zs * (-ad[i]);
      g.d[i + 1] += sgn * zs * (-ad[i + 1] * cd + ad[i + 2] * sd);
      g.d[i + 2] += sgn * zs * (-ad[i + 1] * sd - ad[i + 2] * cd);
    }
    return;
  }

  // part B (surface-deformation) gradients
  const double ir5 = ir3 * ir2;
  const double rd = r + db;
  const double ird = 1.0 / rd;
  const double d11 = ir * ird;
  const double aj2 = xi * yb * ird * d11;
  const double aj5 = -(db + yb * yb * ird) * d11;
  double ak1, ak3, aj3, aj6;
  if (cd != 0.0) {
    const double icd = 1.0 / cd;
    ak1 = xi * (d11 - y11 * sd) * icd;
    ak3 = (q * y11 - yb * d11) * icd;
    aj3 = (ak1 - aj2 * sd) * icd;
    aj6 = (ak3 - aj5 * sd) * icd;
  } else {
    const double ird2 = ird * ird;
    ak1 = xi * q * ird * d11;
    ak3 = sd * ird * (xi2 * d11 - 1.0);
    aj3 = -xi * ird2 * (q2 * d11 - 0.5);
    aj6 = -yb * ird2 * (xi2 * d11 - 0.5);
  }
  const double ak2 = ir + ak3 * sd;
  const double ak4 = xy * cd - ak1 * sd;
  const double aj1 = aj5 * cd - aj6 * sd;
  const double aj4 = -xy - aj2 * cd + aj3 * sd;
  const double t1 = m.a3 * sd, t2 = m.a3 * sd * cd;
  // A + B combined
  as[0] += xi2 * q * y32 - t1 * aj1;
  as[1] += xi * qr3 - t1 * aj2;
  as[2] += -xi * q2 * y32 - t1 * aj3;
  as[3] += -xi * fy - db * x11 + t1 * (xy + aj4);
  as[4] += -ey + t1 * (ir + aj5);
  as[5] += q * fy - t1 * (qy - aj6);
  as[6] += -xi * fz - yb * x11 + t1 * ak1;
  as[7] += -ez + t1 * yb * d11;
  as[8] += q * fz + t1 * ak2;
  ad[0] += xi * qr3 + t2 * aj4;
  ad[1] += et * qr3 + qy + t2 * aj5;
  ad[2] += -q2 * ir3 + t2 * aj6;
  ad[3] += -ey + t2 * aj1;
  ad[4] += -et * gy - xy * sd + t2 * aj2;
  ad[5] += q * gy + t2 * aj3;
  ad[6] += -ez - t2 * ak3;
  ad[7] += -et * gz - xy * cd - t2 * xi * d11;
  ad[8] += q * gz - t2 * ak4;

  // part C (depth) terms: gradients and the displacement part
  const double c = db + z;
  const double x11c = x11 * x11 * x11, y11c = y11 * y11 * y11;
  const double x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * xi2) * x11c * ir2;
  const double y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * et2) * y11c * ir2;
  const double h = q * cd - z;
  const double z32 = sd * ir3 - h * y32;
  const double z53 = 3.0 * sd * ir5 - h * y53;
  const double y0 = y11 - xi2 * y32;
  const double z0 = z32 - xi2 * z53;
  const double ppy = cd * ir3 + q * y32 * sd;
  const double ppz = sd * ir3 - q * y32 * cd;
  const double qq = z * y32 + z32 + z0;
  const double qqy = 3.0 * c * db * ir5 - qq * sd;
  const double qqz = 3.0 * c * yb * ir5 - qq * cd + q * y32;
  const double qr = 3.0 * q * ir5;
  const double cdr = (c + db) * ir3;
  const double yy0 = yb * ir3 - y0 * cd;
  const double a4 = m.a4, a5 = m.a5, a5c = m.a5 * c;
  double cs[12], cq[12];
  cs[0] = a4 * xy * cd - a5 * q * xi * z32;
  cs[1] = a4 * (cd * ir + 2.0 * qy * sd) - a5 * q * c * ir3;
  cs[2] = a4 * qy * cd - a5 * (c * et * ir3 - z * y11 + xi2 * z32);
  cs[3] = a4 * y0 * cd - a5 * q * z0;
  cs[4] = -a4 * xi * (cd * ir3 + 2.0 * q * y32 * sd) + a5 * c * xi * qr;
  cs[5] = -a4 * xi * q * y32 * cd + a5 * xi * (3.0 * c * et * ir5 - qq);
  cs[6] = -a4 * xi * ppy * cd - a5 * xi * qqy;
  cs[7] = a4 * 2.0 * (db * ir3 - y0 * sd) * sd - yb * ir3 * cd -
          a5 * (cdr * sd - et * ir3 - c * yb * qr);
  cs[8] = -a4 * q * ir3 + yy0 * sd +
          a5 * (cdr * cd + c * db * qr - (y0 * cd + q * z0) * sd);
  cs[9] = a4 * xi * ppz * cd - a5 * xi * qqz;
  cs[10] = a4 * 2.0 * (yb * ir3 - y0 * cd) * sd + db * ir3 * cd -
           a5 * (cdr * cd + c * db * qr);
  cs[11] = yy0 * cd - a5 * (cdr * sd - c * yb * qr - y0 * sd * sd + q * z0 * cd);
  cq[0] = a4 * cd * ir - qy * sd - a5c * q * ir3;
  cq[1] = a4 * yb * x11 - a5c * et * q * x32;
  cq[2] = -db * x11 - xy * sd - a5c * (x11 - q2 * x32);
  cq[3] = -a4 * xi * ir3 * cd + a5c * xi * qr + xi * q * y32 * sd;
  cq[4] = -a4 * yb * ir3 + a5c * et * qr;
  cq[5] = db * ir3 - y0 * sd + a5c * ir3 * (1.0 - 3.0 * q2 * ir2);
  cq[6] = -a4 * et * ir3 + y0 * sd * sd - a5 * (cdr * sd - c * yb * qr);
  cq[7] = a4 * (x11 - yb * yb * x32) -
          a5c * ((db + 2.0 * q * cd) * x32 - yb * et * q * x53);
  cq[8] = xi * ppy * sd + yb * db * x32 +
          a5c * ((yb + 2.0 * q * sd) * x32 - yb * q2 * x53);
  cq[9] = -q * ir3 + y0 * sd * cd - a5 * (cdr * cd + c * db * qr);
  cq[10] = a4 * yb * db * x32 -
           a5c * ((yb - 2.0 * q * sd) * x32 + db * et * q * x53);
  cq[11] = -xi * ppz * sd + x11 - db * db * x32 -
           a5c * ((db - 2.0 * q * cd) * x32 - db * q2 * x53);

  // image frame: A + B + z C, plus d(z u_C)/dz = u_C on the z group
#pragma unroll
  for (int grp = 0; grp < 3; ++grp) {
    const int i = 3 * grp, k = i + 3;
    {
      const double zc1 = z * cs[k + 1], zc2 = z * cs[k + 2];
      double d0 = as[i] + z * cs[k];
      double d1 = (as[i + 1] + zc1) * cd - (as[i + 2] + zc2) * sd;
      double d2 = (as[i + 1] - zc1) * sd + (as[i + 2] - zc2) * cd;
      if (grp == 2) {
        d0 += cs[0];
        d1 += cs[1] * cd - cs[2] * sd;
        d2 -= cs[1] * sd + cs[2] * cd;
      }
      g.s[i] += sgn * d0;
      g.s[i + 1] += sgn * d1;
      g.s[i + 2] += sgn * d2;
    }
    {
      const double zc1 = z * cq[k + 1], zc2 = z * cq[k + 2];
      double d0 = ad[i] + z * cq[k];
      double d1 = (ad[i + 1] + zc1) * cd - (ad[i + 2] + zc2) * sd;
      double d2 = (ad[i + 1] - zc1) * sd + (ad[i + 2] - zc2) * cd;
      if (grp == 2) {
        d0 += cq[0];
        d1 += cq[1] * cd - cq[2] * sd;
        d2 -= cq[1] * sd + cq[2] * cd;
      }
      g.d[i] += sgn * d0;
      g.d[i + 1] += sgn * d1;
      g.d[i + 2] += sgn * d2;
    }
  }
}

// Displacement gradients at (x, y, z) in the source's strike frame for a
// rectangle [-L, L] x [-W, W] whose reference point is at depth c (> 0),
// dip (sd, cd) already snapped.  Returns true when the point is on an
// edge (all gradients zero, reference iret = 1).
__device__ __forceinline__ bool okada_grad(double x, double y, double z,
                                           double c, double sd, double cd,
                                           double L, double W, bool full,
                                           const OkMedium& m, OkGrad& g) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    g.s[i] = 0.0;
    g.d[i] = 0.0;
  }
  const double xi0 = ok_snap(x + L), xi1 = ok_snap(x - L);
  for (int img = 0; img < (full ? 1 : 2); ++img) {
    const double d = img ? c - z : c + z;
    const double p = y * cd + d * sd;
    const double q = ok_snap(y * sd - d * cd);
    const double et0 = ok_snap(p + W), et1 = ok_snap(p - W);
    if (q == 0.0 && ((xi0 * xi1 <= 0.0 && et0 * et1 == 0.0) ||
                     (et0 * et1 <= 0.0 && xi0 * xi1 == 0.0))) {
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        g.s[i] = 0.0;
        g.d[i] = 0.0;
      }
      return true;
    }
    const double q2 = q * q;
    const double r12 = sqrt(xi0 * xi0 + et1 * et1 + q2);
    const double r21 = sqrt(xi1 * xi1 + et0 * et0 + q2);
    const double r22 = sqrt(xi1 * xi1 + et1 * et1 + q2);
    const bool kxi0 = (xi0 < 0.0) && (r21 + xi1 < kOkEps);
    const bool kxi1 = (xi0 < 0.0) && (r22 + xi1 < kOkEps);
    const bool ket0 = (et0 < 0.0) && (r12 + et1 < kOkEps);
    const bool ket1 = (et0 < 0.0) && (r22 + et1 < kOkEps);
    // corners (xi_j, et_k): sign + for j == k, - otherwise; the xi-edge
    // flag belongs to the eta index and the eta-edge flag to the xi index
    if (img) {
      ok_corner<true>(xi0, et0, q, kxi0, ket0, sd, cd, z, m, 1.0, g);
      ok_corner<true>(xi1, et0, q, kxi0, ket1, sd, cd, z, m, -1.0, g);
      ok_corner<true>(xi0, et1, q, kxi1, ket0, sd, cd, z, m, -1.0, g);
      ok_corner<true>(xi1, et1, q, kxi1, ket1, sd, cd, z, m, 1.0, g);
    } else {
      ok_corner<false>(xi0, et0, q, kxi0, ket0, sd, cd, z, m, 1.0, g);
      ok_corner<false>(xi1, et0, q, kxi0, ket1, sd, cd, z, m, -1.0, g);
      ok_corner<false>(xi0, et1, q, kxi1, ket0, sd, cd, z, m, -1.0, g);
      ok_corner<false>(xi1, et1, q, kxi1, ket1, sd, cd, z, m, 1.0, g);
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    g.s[i] *= kInv2Pi;
    g.d[i] *= kInv2Pi;
  }
  return false;
}

}  // namespace rsfg
