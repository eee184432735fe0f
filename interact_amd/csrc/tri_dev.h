// Half-space triangular dislocation (Nikkhoo & Walter 2015, GJI 201:1117)
// for operator assembly: stress at a receiver point from unit strike slip
// and unit dip slip on a source triangle, both at once.
//
// Replaces eval_triangle_nw -> tdstresshs (src/green/eval_triangle_nw.c:
// 17-96, src/green/tdstresshs.f90:9-148).  The field is the main
// dislocation in a full space, its image (vertices z -> -z) and the
// harmonic free-surface correction along the three edges.  Strains are the
// symmetric parts of displacement gradients derived symbolically from the
// published displacement fields (nw_grad_gen.h): per angular dislocation
// the gradients for the three unit Burgers vectors share one evaluation,
// and the two slip directions are linear combinations of them.  The
// triangle's Burgers-function gradient completes the full-space strain.
//
// Frames follow the reference (get_tdcs_base_vectors, src/green/
// tddisphs.f90:790-817): TDCS axes (normal, strike, dip), origin at the
// second vertex, slip (strike, dip) = Burgers vector (0, 1, 0) / (0, 0, 1).
#pragma once
#include <hip/hip_runtime.h>

#define NW_FN static __device__ __forceinline__
#include "nw_grad_gen.h"

namespace rsfg {

// the reference's pi in the angle offsets and its vertical-edge tolerance
// (src/green/tdstresshs.f90 pi parameter, EPS_FOR_FORTRAN): an edge whose
// angle beta from the downward vertical is within the tolerance of 0 or pi
// carries no free-surface correction.  Edges a little further from vertical
// make the correction's two end terms grow like cot(beta)^2 and cancel, so
// a caller may widen the tolerance (rsfg_set_triangle_tolerance).
constexpr double kTriPi = 3.14159265358979;
constexpr double kTriEps = 5.0e-15;
constexpr double kTriNu = 0.25;

__device__ __forceinline__ double tdot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void tunit(double* a) {
  const double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a[0] /= l;
  a[1] /= l;
  a[2] /= l;
}

// TDCS base vectors of the triangle (normal, strike, dip)
__device__ __forceinline__ void tri_frame(const double* p1, const double* p2, const double* p3,
                                          double* vn, double* vs, double* vd) {
  const double a[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double b[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  vn[0] = a[1] * b[2] - a[2] * b[1];
  vn[1] = a[2] * b[0] - a[0] * b[2];
  vn[2] = a[0] * b[1] - a[1] * b[0];
  tunit(vn);
  vs[0] = -vn[1];
  vs[1] = vn[0];
  vs[2] = 0.0;
  if (sqrt(tdot(vs, vs)) == 0.0) {
    vs[0] = 0.0;
    vs[1] = (p1[2] > 0.0) ? -vn[2] : vn[2];
  }
  tunit(vs);
  vd[0] = vn[1] * vs[2] - vn[2] * vs[1];
  vd[1] = vn[2] * vs[0] - vn[0] * vs[2];
  vd[2] = vn[0] * vs[1] - vn[1] * vs[0];
}

// angular-dislocation arrangement of a point relative to the triangle
// (trimodefinder, src/green/tddisphs.f90:331-364); 0 = on an edge line in
// the triangle's plane (the reference's NaN stress)
__device__ __forceinline__ int tri_mode(double x, double y, double z, double p1y, double p1z,
                                        double p3y, double p3z) {
  // second vertex at the origin
  const double den = (0.0 - p3z) * (p1y - p3y) + (p3y - 0.0) * (p1z - p3z);
  const double a = ((0.0 - p3z) * (x - p3y) + (p3y - 0.0) * (y - p3z)) / den;
  const double b = ((p3z - p1z) * (x - p3y) + (p1y - p3y) * (y - p3z)) / den;
  const double c = 1.0 - a - b;
  int m = 1;
  if (a <= 0.0 && b > c && c > a) m = -1;
  if (b <= 0.0 && c > a && a > b) m = -1;
  if (c <= 0.0 && a > b && b > c) m = -1;
  if (a == 0.0 && b >= 0.0 && c >= 0.0) m = 0;
  if (a >= 0.0 && b == 0.0 && c >= 0.0) m = 0;
  if (a >= 0.0 && b >= 0.0 && c == 0.0) m = 0;
  if (m == 0 && z != 0.0) m = 1;
  return m;
}

// Every contribution below enters the result only through the traction it
// exerts on the receiver plane, t = S n with S = 2 mu E + lambda tr(E) I:
// t = 2 mu (E n) + lambda tr(E) n, and E n of a strain given in a rotated
// frame R (E = R E' R^T) is R (E' (R^T n)).  So each contribution is
// projected as soon as it is formed: the receiver normal goes into the
// contribution's frame, and only E' n' and tr(E') are kept, six numbers
// per pair in all instead of the strain and stress tensors of two slips.
__device__ __forceinline__ void traction_add(const double* en, double trE, const double* n,
                                             double mu2, double lambda, double* t) {
#pragma unroll
  for (int a = 0; a < 3; ++a) t[a] += mu2 * en[a] + lambda * trE * n[a];
}

// Full-space triangle (the main dislocation, or its image): adds the
// receiver-plane tractions of unit strike slip (t[0]) and unit dip slip
// (t[1]) for the receiver normal nr; returns the TDCS strike and dip vectors
// (the global Burgers vectors of the two unit slips).  false: point on an
// edge line in the triangle's plane.
__device__ __forceinline__ bool td_fs_traction(const double* P, const double* p1,
                                               const double* p2, const double* p3,
                                               const double* nr, double mu2, double lambda,
                                               double t[2][3], double* vs, double* vd) {
  double vn[3];
  tri_frame(p1, p2, p3, vn, vs, vd);
  const double dP[3] = {P[0] - p2[0], P[1] - p2[1], P[2] - p2[2]};
  const double d1[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double d3[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
  const double q[3] = {tdot(vn, dP), tdot(vs, dP), tdot(vd, dP)};
  const double t1[3] = {tdot(vn, d1), tdot(vs, d1), tdot(vd, d1)};
  const double t3[3] = {tdot(vn, d3), tdot(vs, d3), tdot(vd, d3)};
  double e12[3] = {-t1[0], -t1[1], -t1[2]};
  double e13[3] = {t3[0] - t1[0], t3[1] - t1[1], t3[2] - t1[2]};
  double e23[3] = {t3[0], t3[1], t3[2]};
  tunit(e12);
  tunit(e13);
  tunit(e23);
  const int m = tri_mode(q[1], q[2], q[0], t1[1], t1[2], t3[1], t3[2]);
  if (m == 0) return false;
  const double ang[3] = {acos(tdot(e12, e13)), acos(-tdot(e12, e23)), acos(tdot(e23, e13))};
  const double sgn = (m == 1) ? 1.0 : -1.0;
  // edge k: apex (y, z) and direction (y, z components) of its side vector
  const double apy[3] = {t1[1], 0.0, t3[1]}, apz[3] = {t1[2], 0.0, t3[2]};
  const double svy[3] = {-sgn * e13[1], sgn * e12[1], sgn * e23[1]};
  const double svz[3] = {-sgn * e13[2], sgn * e12[2], sgn * e23[2]};
  // the receiver normal in the TDCS frame (columns vn, vs, vd)
  const double nt[3] = {tdot(vn, nr), tdot(vs, nr), tdot(vd, nr)};
  double tl[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};   // E n in TDCS
  double tr[2] = {0.0, 0.0};                              // tr E
  for (int k = 0; k < 3; ++k) {
    // angular-dislocation frame: rotation about the normal that takes the
    // side vector to -z (get_tdcd_adcs, src/green/tddisphs.f90:826-842);
    // TDCS strain = B E_a B^T with B = [[1,0,0],[0,c,s],[0,-s,c]]
    const double c = svz[k], s = svy[k];
    const double dy = q[1] - apy[k], dz = q[2] - apz[k];
    double sA, cA;
    sincos(-kTriPi + ang[k], &sA, &cA);
    double gb[3][3][3];
    nw_adis_gradb(q[0], c * dy - s * dz, s * dy + c * dz, cA, sA, kTriNu, gb);
    const double nb[3] = {nt[0], c * nt[1] - s * nt[2], s * nt[1] + c * nt[2]};   // B^T nt
    // unit strike slip: b = (0, c, s) in the dislocation frame; dip: (0, -s, c)
    const double wy[2] = {c, -s}, wz[2] = {s, c};
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      // E_a = sym(G), G = wy gb[1] + wz gb[2]: E_a nb = (G nb + G^T nb) / 2
      double u[3], te = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double gi = 0.0, gti = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          gi += (wy[sl] * gb[1][i][j] + wz[sl] * gb[2][i][j]) * nb[j];
          gti += (wy[sl] * gb[1][j][i] + wz[sl] * gb[2][j][i]) * nb[j];
        }
        u[i] = 0.5 * (gi + gti);
        te += wy[sl] * gb[1][i][i] + wz[sl] * gb[2][i][i];
      }
      tl[sl][0] += u[0];
      tl[sl][1] += c * u[1] + s * u[2];
      tl[sl][2] += -s * u[1] + c * u[2];
      tr[sl] += te;
    }
  }
  // Burgers function of the triangle: b = e_strike (TDCS y) or e_dip (TDCS
  // z); its strain is (e_c g^T + g e_c^T) / 2 with g = gf[0]
  {
    double gf[3][3];
    nw_burg_grad(q[0], q[1], q[2], t1[1], t1[2], t3[1], t3[2], gf);
    const double gn = gf[0][0] * nt[0] + gf[0][1] * nt[1] + gf[0][2] * nt[2];
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      const int c = 1 + sl;
      tl[sl][c] += 0.5 * gn;
#pragma unroll
      for (int j = 0; j < 3; ++j) tl[sl][j] += 0.5 * gf[0][j] * nt[c];
      tr[sl] += gf[0][c];
    }
  }
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    double en[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) en[a] = vn[a] * tl[sl][0] + vs[a] * tl[sl][1] + vd[a] * tl[sl][2];
    traction_add(en, tr[sl], nr, mu2, lambda, t[sl]);
  }
  return true;
}

// free-surface correction on edge PA-PB for the two global Burgers vectors
// bs (strike slip) and bd (dip slip), its receiver-plane tractions added to
// t (AngSetupFSC_S, src/green/tdstresshs.f90:371-476).  The edge frame A is
// symmetric (its own transpose); the two ends' gradients are consumed one
// after the other.  vtol: the vertical-edge tolerance (kTriEps is the
// reference's).
__device__ __forceinline__ void fsc_edge_traction(const double* P, const double* bs,
                                                  const double* bd, const double* pa,
                                                  const double* pb, const double* nr,
                                                  double mu2, double lambda, double vtol,
                                                  double t[2][3]) {
  const double sv[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
  const double beta = acos(-sv[2] / sqrt(tdot(sv, sv)));
  if (fabs(beta) < vtol || fabs(kTriPi - beta) < vtol) return;
  double e1[3] = {sv[0], sv[1], 0.0};
  tunit(e1);
  // frame (e1, (-e_z) x e1, -e_z)
  const double A[3][3] = {{e1[0], e1[1], 0.0}, {e1[1], -e1[0], 0.0}, {0.0, 0.0, -1.0}};
  const double d[3] = {P[0] - pa[0], P[1] - pa[1], P[2] - pa[2]};
  double yA[3], yB[3], bbs[3], bbd[3], ne[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    yA[i] = A[i][0] * d[0] + A[i][1] * d[1] + A[i][2] * d[2];
    yB[i] = yA[i] - (A[i][0] * sv[0] + A[i][1] * sv[1] + A[i][2] * sv[2]);
    bbs[i] = A[i][0] * bs[0] + A[i][1] * bs[1] + A[i][2] * bs[2];
    bbd[i] = A[i][0] * bd[0] + A[i][1] * bd[1] + A[i][2] * bd[2];
    ne[i] = A[i][0] * nr[0] + A[i][1] * nr[1] + A[i][2] * nr[2];
  }
  double sB, cB;
  sincos((beta * yA[0] >= 0.0) ? -kTriPi + beta : beta, &sB, &cB);
  // V = sym(sum_c bb[c] (gB - gA)[c]) in the edge frame: V ne and tr V
  double u[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, te[2] = {0.0, 0.0};
#pragma unroll
  for (int end = 0; end < 2; ++end) {
    const double sg = end ? 1.0 : -1.0;
    const double* y = end ? yB : yA;
    double g[3][3][3];
    nw_fsc_gradb(y[0], y[1], y[2], cB, sB, kTriNu, end ? -pb[2] : -pa[2], g);
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      const double* bb = sl ? bbd : bbs;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double gi = 0.0, gti = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            gi += bb[c] * g[c][i][j] * ne[j];
            gti += bb[c] * g[c][j][i] * ne[j];
          }
        u[sl][i] += sg * 0.5 * (gi + gti);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) te[sl] += sg * bb[c] * (g[c][0][0] + g[c][1][1] + g[c][2][2]);
    }
  }
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    double en[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) en[a] = A[a][0] * u[sl][0] + A[a][1] * u[sl][1] + A[a][2] * u[sl][2];
    traction_add(en, te[sl], nr, mu2, lambda, t[sl]);
  }
}

// Traction t = S n on the receiver plane (normal nr) for unit strike slip
// (t[0]) and unit dip slip (t[1]) on the triangle v[0..8], internal elastic
// constants (mu2 = 2 mu, lambda): the main dislocation, the free-surface
// correction of its three edges, and the image dislocation.  false when the
// result is not finite (the reference's iret = 1, entry 0).  A triangle with
// all vertices at z = 0 gives zero.  vtol: the vertical-edge tolerance.
__device__ __forceinline__ bool tri_unit_traction(const double* P, const double* v,
                                                  const double* nr, double mu2, double lambda,
                                                  double vtol, double t[2][3]) {
#pragma unroll
  for (int s = 0; s < 2; ++s) t[s][0] = t[s][1] = t[s][2] = 0.0;
  if (v[2] == 0.0 && v[5] == 0.0 && v[8] == 0.0) return true;
  double vs[3], vd[3];
  if (!td_fs_traction(P, v, v + 3, v + 6, nr, mu2, lambda, t, vs, vd)) return false;
  fsc_edge_traction(P, vs, vd, v, v + 3, nr, mu2, lambda, vtol, t);
  fsc_edge_traction(P, vs, vd, v + 3, v + 6, nr, mu2, lambda, vtol, t);
  fsc_edge_traction(P, vs, vd, v + 6, v, nr, mu2, lambda, vtol, t);
  const double w[9] = {v[0], v[1], -v[2], v[3], v[4], -v[5], v[6], v[7], -v[8]};
  double vsi[3], vdi[3];
  if (!td_fs_traction(P, w, w + 3, w + 6, nr, mu2, lambda, t, vsi, vdi)) return false;
  bool ok = true;
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int a = 0; a < 3; ++a) ok = ok && isfinite(t[s][a]);
  return ok;
}

}  // namespace rsfg
