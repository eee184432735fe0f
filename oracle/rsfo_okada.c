/*
  Oracle restatement of the Okada (1992) rectangular-dislocation solution
  (BSSA 82:1018-1040, "Internal deformation due to shear and tensile faults
  in a half-space") and of the reference's frame handling around it.

  TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).

  Structure follows the paper: the half-space field is
      u = u_A(real source) [sign-flipped frame]  +  [u_A + u_B + z u_C](image)
  summed over the four rectangle corners with the Chinnery sign pattern;
  the reference's numerical guards are kept as behaviour:
   - coordinates closer than OKD_EPS to 0 snap to 0 (src/green/dc3d.F:646-660,
     :1274-1276), |cos(dip)| < OKD_EPS snaps to 0 (:1186-1190);
   - an observation point on a rectangle edge is singular: all outputs 0,
     iret = 1 (:665-670, :726-731, :795-813);
   - points on the negative extension of an edge use the log(R - xi) limit
     forms (KXI/KET, :672-685, :1293-1312);
   - full space keeps the real-source term only (:712-715).
*/
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "rsfo.h"

/* The scalar type of the per-pair arithmetic.  double for the oracle; the
   operation-count harness (opcount.cpp) compiles this file as C++ with a
   counting type, which must not change what is computed: OKD_TO_DOUBLE
   reads a value back and OKD_NOTE_SINCOS records the one sin/cos pair of
   the dip that DC3D evaluates per call. */
#ifndef OKD_REAL
#define OKD_REAL double
#endif
#ifndef OKD_TO_DOUBLE
#define OKD_TO_DOUBLE(v) (v)
#endif
#ifndef OKD_NOTE_SINCOS
#define OKD_NOTE_SINCOS()
#endif
typedef OKD_REAL real;

#define OKD_EPS 1.0e-8                        /* EPS_DC3D_PREC, dc3d.F:515 */
#define OKD_PI2 6.283185307179586477          /* 2 pi, dc3d.F:836 */
#define OKD_DEG2RAD 0.0174532925199432957692369076848861271344287188854172545609719144

typedef struct {
  real alp1, alp2, alp3, alp4, alp5;
  real sd, cd, sdsd, cdcd, sdcd;
} okd_const;

/* medium and dip constants (DCCON0) */
static void okd_const_init(okd_const *k, real alpha, double dip_deg)
{
  double s, c;
  k->alp1 = (1.0 - alpha) / 2.0;
  k->alp2 = alpha / 2.0;
  k->alp3 = (1.0 - alpha) / alpha;
  k->alp4 = 1.0 - alpha;
  k->alp5 = alpha;
  sincos(dip_deg * OKD_DEG2RAD, &s, &c);    /* my_sincos_deg_ftnd, src/util/mysincos.c:28-33 */
  OKD_NOTE_SINCOS();
  k->sd = s;
  k->cd = c;
  if (fabs(c) < OKD_EPS) {
    k->cd = 0.0;
    if (s > 0.0) k->sd = 1.0;
    if (s < 0.0) k->sd = -1.0;
  }
  k->sdsd = k->sd * k->sd;
  k->cdcd = k->cd * k->cd;
  k->sdcd = k->sd * k->cd;
}

/* corner geometry of the finite source (DCCON2) */
typedef struct {
  real xi, et, q, xi2, et2, q2, r, r2, r3, r5;
  real yb, db;                 /* y-bar, d-bar of the paper */
  real tt, alx, ale, x11, y11, x32, y32;
  real ey, ez, fy, fz, gy, gz;
} okd_corner;

static real snap(real v) { return (fabs(v) < OKD_EPS) ? 0.0 : v; }

static void okd_corner_init(okd_corner *g, real xi, real et, real q,
                            const okd_const *k, int kxi, int ket)
{
  real t;
  g->xi = snap(xi);
  g->et = snap(et);
  g->q = snap(q);
  g->xi2 = g->xi * g->xi;
  g->et2 = g->et * g->et;
  g->q2 = g->q * g->q;
  g->r2 = g->xi2 + g->et2 + g->q2;
  g->r = sqrt(g->r2);
  /* r == 0 cannot reach here: the corner-at-observation case is rejected
     as an edge singularity before any corner is evaluated */
  g->r3 = g->r * g->r2;
  g->r5 = g->r3 * g->r2;
  g->yb = g->et * k->cd + g->q * k->sd;
  g->db = g->et * k->sd - g->q * k->cd;
  g->tt = (g->q == 0.0) ? 0.0 : atan(g->xi * g->et / (g->q * g->r));
  if (kxi) {
    g->alx = -log(g->r - g->xi);
    g->x11 = 0.0;
    g->x32 = 0.0;
  } else {
    real rxi = g->r + g->xi;
    g->alx = log(rxi);
    g->x11 = 1.0 / (g->r * rxi);
    g->x32 = (g->r + rxi) * g->x11 * g->x11 / g->r;
  }
  if (ket) {
    g->ale = -log(g->r - g->et);
    g->y11 = 0.0;
    g->y32 = 0.0;
  } else {
    real ret = g->r + g->et;
    g->ale = log(ret);
    g->y11 = 1.0 / (g->r * ret);
    g->y32 = (g->r + ret) * g->y11 * g->y11 / g->r;
  }
  t = g->q / g->r3;
  g->ey = k->sd / g->r - g->yb * t;
  g->ez = k->cd / g->r + g->db * t;
  g->fy = g->db / g->r3 + g->xi2 * g->y32 * k->sd;
  g->fz = g->yb / g->r3 + g->xi2 * g->y32 * k->cd;
  g->gy = 2.0 * g->x11 * k->sd - g->yb * g->q * g->x32;
  g->gz = 2.0 * g->x11 * k->cd + g->db * g->q * g->x32;
}

/* accumulate s * du into u */
static void axpy12(real *u, real s, const real *du)
{
  int i;
  for (i = 0; i < 12; i++) u[i] += s * du[i];
}

/* Part A: infinite-medium terms (UA) */
static void okd_part_a(const okd_corner *g, const okd_const *k, real d1,
                       real d2, real *u)
{
  real du[12];
  const real xi = g->xi, et = g->et, q = g->q, r = g->r, r3 = g->r3;
  const real xy = xi * g->y11, qx = q * g->x11, qy = q * g->y11;
  memset(u, 0, 12 * sizeof(real));
  if (d1 != 0.0) {
    du[0] = g->tt / 2.0 + k->alp2 * xi * qy;
    du[1] = k->alp2 * q / r;
    du[2] = k->alp1 * g->ale - k->alp2 * q * qy;
    du[3] = -k->alp1 * qy - k->alp2 * q * g->xi2 * g->y32;
    du[4] = -k->alp2 * q * xi / r3;
    du[5] = k->alp1 * xy + k->alp2 * xi * g->q2 * g->y32;
    du[6] = k->alp1 * xy * k->sd + k->alp2 * xi * g->fy + g->db / 2.0 * g->x11;
    du[7] = k->alp2 * g->ey;
    du[8] = k->alp1 * (k->cd / r + qy * k->sd) - k->alp2 * q * g->fy;
    du[9] = k->alp1 * xy * k->cd + k->alp2 * xi * g->fz + g->yb / 2.0 * g->x11;
    du[10] = k->alp2 * g->ez;
    du[11] = -k->alp1 * (k->sd / r - qy * k->cd) - k->alp2 * q * g->fz;
    axpy12(u, d1 / OKD_PI2, du);
  }
  if (d2 != 0.0) {
    du[0] = k->alp2 * q / r;
    du[1] = g->tt / 2.0 + k->alp2 * et * qx;
    du[2] = k->alp1 * g->alx - k->alp2 * q * qx;
    du[3] = -k->alp2 * q * xi / r3;
    du[4] = -qy / 2.0 - k->alp2 * q * et / r3;
    du[5] = k->alp1 / r + k->alp2 * g->q2 / r3;
    du[6] = k->alp2 * g->ey;
    du[7] = k->alp1 * g->x11 * g->db + xy / 2.0 * k->sd + k->alp2 * et * g->gy;
    du[8] = k->alp1 * g->x11 * g->yb - k->alp2 * q * g->gy;
    du[9] = k->alp2 * g->ez;
    du[10] = k->alp1 * g->x11 * g->yb + xy / 2.0 * k->cd + k->alp2 * et * g->gz;
    du[11] = -k->alp1 * g->x11 * g->db - k->alp2 * q * g->gz;
    axpy12(u, d2 / OKD_PI2, du);
  }
}

/* Part B: surface-deformation related terms (UB) */
static void okd_part_b(const okd_corner *g, const okd_const *k, real d1,
                       real d2, real *u)
{
  real du[12];
  const real xi = g->xi, et = g->et, q = g->q, r = g->r, r3 = g->r3;
  const real yb = g->yb, db = g->db;
  real rd = r + db, d11 = 1.0 / (r * rd);
  real aj2 = xi * yb / rd * d11, aj5 = -(db + yb * yb / rd) * d11;
  real ai1, ai2, ai3, ai4, ak1, ak2, ak3, ak4, aj1, aj3, aj4, aj6;
  real xy = xi * g->y11, qx = q * g->x11, qy = q * g->y11, t;
  if (k->cd != 0.0) {
    if (xi == 0.0) {
      ai4 = 0.0;
    } else {
      real xx = sqrt(g->xi2 + g->q2);
      ai4 = 1.0 / k->cdcd * (xi / rd * k->sdcd +
            2.0 * atan((et * (xx + q * k->cd) + xx * (r + xx) * k->sd) /
                       (xi * (r + xx) * k->cd)));
    }
    ai3 = (yb * k->cd / rd - g->ale + k->sd * log(rd)) / k->cdcd;
    ak1 = xi * (d11 - g->y11 * k->sd) / k->cd;
    ak3 = (q * g->y11 - yb * d11) / k->cd;
    aj3 = (ak1 - aj2 * k->sd) / k->cd;
    aj6 = (ak3 - aj5 * k->sd) / k->cd;
  } else {
    real rd2 = rd * rd;
    ai3 = (et / rd + yb * q / rd2 - g->ale) / 2.0;
    ai4 = xi * yb / rd2 / 2.0;
    ak1 = xi * q / rd * d11;
    ak3 = k->sd / rd * (g->xi2 * d11 - 1.0);
    aj3 = -xi / rd2 * (g->q2 * d11 - 0.5);
    aj6 = -yb / rd2 * (g->xi2 * d11 - 0.5);
  }
  ai1 = -xi / rd * k->cd - ai4 * k->sd;
  ai2 = log(rd) + ai3 * k->sd;
  ak2 = 1.0 / r + ak3 * k->sd;
  ak4 = xy * k->cd - ak1 * k->sd;
  aj1 = aj5 * k->cd - aj6 * k->sd;
  aj4 = -xy - aj2 * k->cd + aj3 * k->sd;
  memset(u, 0, 12 * sizeof(real));
  if (d1 != 0.0) {
    t = k->alp3 * k->sd;
    du[0] = -xi * qy - g->tt - t * ai1;
    du[1] = -q / r + t * yb / rd;
    du[2] = q * qy - t * ai2;
    du[3] = g->xi2 * q * g->y32 - t * aj1;
    du[4] = xi * q / r3 - t * aj2;
    du[5] = -xi * g->q2 * g->y32 - t * aj3;
    du[6] = -xi * g->fy - db * g->x11 + t * (xy + aj4);
    du[7] = -g->ey + t * (1.0 / r + aj5);
    du[8] = q * g->fy - t * (qy - aj6);
    du[9] = -xi * g->fz - yb * g->x11 + t * ak1;
    du[10] = -g->ez + t * yb * d11;
    du[11] = q * g->fz + t * ak2;
    axpy12(u, d1 / OKD_PI2, du);
  }
  if (d2 != 0.0) {
    t = k->alp3 * k->sdcd;
    du[0] = -q / r + t * ai3;
    du[1] = -et * qx - g->tt - t * xi / rd;
    du[2] = q * qx + t * ai4;
    du[3] = xi * q / r3 + t * aj4;
    du[4] = et * q / r3 + qy + t * aj5;
    du[5] = -g->q2 / r3 + t * aj6;
    du[6] = -g->ey + t * aj1;
    du[7] = -et * g->gy - xy * k->sd + t * aj2;
    du[8] = q * g->gy + t * aj3;
    du[9] = -g->ez - t * ak3;
    du[10] = -et * g->gz - xy * k->cd - t * xi * d11;
    du[11] = q * g->gz - t * ak4;
    axpy12(u, d2 / OKD_PI2, du);
  }
}

/* Part C: depth-dependent terms (UC) */
static void okd_part_c(const okd_corner *g, const okd_const *k, real z,
                       real d1, real d2, real *u)
{
  real du[12];
  const real xi = g->xi, et = g->et, q = g->q, r = g->r, r2 = g->r2;
  const real r3 = g->r3, r5 = g->r5, yb = g->yb, db = g->db;
  const real sd = k->sd, cd = k->cd;
  real c = db + z;
  real x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * g->xi2) * g->x11 * g->x11 * g->x11 / r2;
  real y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * g->et2) * g->y11 * g->y11 * g->y11 / r2;
  real h = q * cd - z;
  real z32 = sd / r3 - h * g->y32;
  real z53 = 3.0 * sd / r5 - h * y53;
  real y0 = g->y11 - g->xi2 * g->y32;
  real z0 = z32 - g->xi2 * z53;
  real ppy = cd / r3 + q * g->y32 * sd;
  real ppz = sd / r3 - q * g->y32 * cd;
  real qq = z * g->y32 + z32 + z0;
  real qqy = 3.0 * c * db / r5 - qq * sd;
  real qqz = 3.0 * c * yb / r5 - qq * cd + q * g->y32;
  real xy = xi * g->y11, qy = q * g->y11;
  real qr = 3.0 * q / r5;
  real cdr = (c + db) / r3;
  real yy0 = yb / r3 - y0 * cd;
  const real a4 = k->alp4, a5 = k->alp5;
  memset(u, 0, 12 * sizeof(real));
  if (d1 != 0.0) {
    du[0] = a4 * xy * cd - a5 * q * xi * z32;
    du[1] = a4 * (cd / r + 2.0 * qy * sd) - a5 * q * c / r3;
    du[2] = a4 * qy * cd - a5 * (c * et / r3 - z * g->y11 + g->xi2 * z32);
    du[3] = a4 * y0 * cd - a5 * q * z0;
    du[4] = -a4 * xi * (cd / r3 + 2.0 * q * g->y32 * sd) + a5 * c * xi * qr;
    du[5] = -a4 * xi * q * g->y32 * cd + a5 * xi * (3.0 * c * et / r5 - qq);
    du[6] = -a4 * xi * ppy * cd - a5 * xi * qqy;
    du[7] = a4 * 2.0 * (db / r3 - y0 * sd) * sd - yb / r3 * cd
          - a5 * (cdr * sd - et / r3 - c * yb * qr);
    du[8] = -a4 * q / r3 + yy0 * sd + a5 * (cdr * cd + c * db * qr - (y0 * cd + q * z0) * sd);
    du[9] = a4 * xi * ppz * cd - a5 * xi * qqz;
    du[10] = a4 * 2.0 * (yb / r3 - y0 * cd) * sd + db / r3 * cd - a5 * (cdr * cd + c * db * qr);
    du[11] = yy0 * cd - a5 * (cdr * sd - c * yb * qr - y0 * k->sdsd + q * z0 * cd);
    axpy12(u, d1 / OKD_PI2, du);
  }
  if (d2 != 0.0) {
    real a5c = a5 * c;
    du[0] = a4 * cd / r - qy * sd - a5c * q / r3;
    du[1] = a4 * yb * g->x11 - a5c * et * q * g->x32;
    du[2] = -db * g->x11 - xy * sd - a5c * (g->x11 - g->q2 * g->x32);
    du[3] = -a4 * xi / r3 * cd + a5c * xi * qr + xi * q * g->y32 * sd;
    du[4] = -a4 * yb / r3 + a5c * et * qr;
    du[5] = db / r3 - y0 * sd + a5c / r3 * (1.0 - 3.0 * g->q2 / r2);
    du[6] = -a4 * et / r3 + y0 * k->sdsd - a5 * (cdr * sd - c * yb * qr);
    du[7] = a4 * (g->x11 - yb * yb * g->x32) - a5c * ((db + 2.0 * q * cd) * g->x32 - yb * et * q * x53);
    du[8] = xi * ppy * sd + yb * db * g->x32 + a5c * ((yb + 2.0 * q * sd) * g->x32 - yb * g->q2 * x53);
    du[9] = -q / r3 + y0 * k->sdcd - a5 * (cdr * cd + c * db * qr);
    du[10] = a4 * yb * db * g->x32 - a5c * ((yb - 2.0 * q * sd) * g->x32 + db * et * q * x53);
    du[11] = -xi * ppz * sd + g->x11 - db * db * g->x32 - a5c * ((db - 2.0 * q * cd) * g->x32 - db * g->q2 * x53);
    axpy12(u, d2 / OKD_PI2, du);
  }
}

/* edge-singularity test of dc3d.F:665-670 */
static int on_edge(real q, const real xi[2], const real et[2])
{
  return (q == 0.0) && (((xi[0] * xi[1] <= 0.0) && (et[0] * et[1] == 0.0)) ||
                        ((et[0] * et[1] <= 0.0) && (xi[0] * xi[1] == 0.0)));
}

/* negative-extension flags, dc3d.F:672-685 */
static void ext_flags(real q, const real xi[2], const real et[2], int kxi[2],
                      int ket[2])
{
  real r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
  real r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
  real r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
  kxi[0] = (xi[0] < 0.0) && (r21 + xi[1] < OKD_EPS);
  kxi[1] = (xi[0] < 0.0) && (r22 + xi[1] < OKD_EPS);
  ket[0] = (et[0] < 0.0) && (r12 + et[1] < OKD_EPS);
  ket[1] = (et[0] < 0.0) && (r22 + et[1] < OKD_EPS);
}

static int okd_dc3d(real alpha, real x, real y, real z, real depth, double dip_deg,
                    real al1, real al2, real aw1, real aw2, real d1, real d2,
                    int full_space, real *u)
{
  okd_const k;
  okd_corner g;
  real xi[2], et[2], q, d, p, ua[12], ub[12], uc[12], du[12];
  int kxi[2], ket[2], i, jx, ke, src;
  memset(u, 0, 12 * sizeof(real));
  okd_const_init(&k, alpha, dip_deg);
  xi[0] = snap(x - al1);
  xi[1] = snap(x - al2);
  /* src 0: real source (depth + z), src 1: image source (depth - z) */
  for (src = 0; src < (full_space ? 1 : 2); src++) {
    d = (src == 0) ? depth + z : depth - z;
    p = y * k.cd + d * k.sd;
    q = snap(y * k.sd - d * k.cd);
    et[0] = snap(p - aw1);
    et[1] = snap(p - aw2);
    if (on_edge(q, xi, et)) {
      memset(u, 0, 12 * sizeof(real));
      return 1;
    }
    ext_flags(q, xi, et, kxi, ket);
    for (ke = 0; ke < 2; ke++) {
      for (jx = 0; jx < 2; jx++) {
        okd_corner_init(&g, xi[jx], et[ke], q, &k, kxi[ke], ket[jx]);
        okd_part_a(&g, &k, d1, d2, ua);
        if (src == 0) {
          /* real source: image-frame reflection of part A (dc3d.F:693-701) */
          for (i = 0; i < 12; i += 3) {
            du[i] = -ua[i];
            du[i + 1] = -ua[i + 1] * k.cd + ua[i + 2] * k.sd;
            du[i + 2] = -ua[i + 1] * k.sd - ua[i + 2] * k.cd;
            if (i == 9) {
              du[i] = -du[i];
              du[i + 1] = -du[i + 1];
              du[i + 2] = -du[i + 2];
            }
          }
        } else {
          okd_part_b(&g, &k, d1, d2, ub);
          okd_part_c(&g, &k, z, d1, d2, uc);
          for (i = 0; i < 12; i += 3) {
            real zc1 = z * uc[i + 1], zc2 = z * uc[i + 2];
            real ab1 = ua[i + 1] + ub[i + 1], ab2 = ua[i + 2] + ub[i + 2];
            du[i] = ua[i] + ub[i] + z * uc[i];
            du[i + 1] = (ab1 + zc1) * k.cd - (ab2 + zc2) * k.sd;
            du[i + 2] = (ab1 - zc1) * k.sd + (ab2 - zc2) * k.cd;
          }
          /* z-derivative of the z u_C product (dc3d.F:774-776) */
          du[9] = du[9] + uc[0];
          du[10] = du[10] + uc[1] * k.cd - uc[2] * k.sd;
          du[11] = du[11] - uc[1] * k.sd - uc[2] * k.cd;
        }
        /* corner sum: + for (xi1,et1),(xi2,et2), - for the mixed ones */
        if (jx + ke == 1) {
          for (i = 0; i < 12; i++) u[i] -= du[i];
        } else {
          for (i = 0; i < 12; i++) u[i] += du[i];
        }
      }
    }
  }
  return 0;
}

int rsfo_dc3d(double alpha, double x, double y, double z, double depth,
              double dip_deg, double al1, double al2, double aw1, double aw2,
              double d1, double d2, int full_space, double u[12])
{
  real ur[12];
  int i, iret = okd_dc3d(alpha, x, y, z, depth, dip_deg, al1, al2, aw1, aw2, d1, d2,
                         full_space, ur);
  for (i = 0; i < 12; i++) u[i] = OKD_TO_DOUBLE(ur[i]);
  return iret;
}

/* ---------------- frames, stress, traction ---------------- */

void rsfo_patch_frame(float strike, float dip, double *fr)
{
  double ca, sa, sdip, cdip, alpha = 90.0 - (double)strike;
  /* my_sincos_degd(alpha) and my_sincos_deg(dip), src/util/mysincos.c:86-91 */
  sincos(alpha * OKD_DEG2RAD, &sa, &ca);
  sincos((double)dip * OKD_DEG2RAD, &sdip, &cdip);
  fr[0] = ca;
  fr[1] = sa;
  /* calc_quad_base_vecs, src/geometry/geometry.c:128-148 */
  fr[2] = sa * sdip;
  fr[3] = -ca * sdip;
  fr[4] = cdip;
  fr[5] = ca;
  fr[6] = sa;
  fr[7] = 0.0;
  fr[8] = -sa * cdip;
  fr[9] = ca * cdip;
  fr[10] = sdip;
  fr[11] = (double)dip;
  fr[12] = 0.0;
}

/* elastic constants at the internal shear modulus 1e4, nu = 0.25
   (calc_medium_elastic_parameters, src/read_geometry.c:702-719) */
#define OKD_MU 1.0e4
#define OKD_NU 0.25

static void okd_traction_pair(const double *xr, const double *frr, const double *xs,
                              const double *frs, double ls, double ws, int full_space,
                              real out[2][3], int iret[2])
{
  const double mu = OKD_MU, mu2 = 2.0 * OKD_MU;
  const double lambda = (mu2 * OKD_NU) / (1.0 - 2.0 * OKD_NU);
  const double alpha = (lambda + mu) / (lambda + mu2);
  real dx, dy, xl, yl, u[12], iso, s[3][3], g[3][3], tr[3];
  const real xr0 = xr[0], xr1 = xr[1], xs0 = xs[0], xs1 = xs[1];
  const real c = frs[0], sn = -frs[1];     /* rotate back with (cos a, -sin a) */
  real cc, ssq, sc;
  int sidx, i, j;
  /* shift, keep z absolute, rotate into the strike frame (eval_okada.c:95-99) */
  dx = xr0 - xs0;
  dy = xr1 - xs1;
  xl = frs[0] * dx + frs[1] * dy;     /* rotate_vec, src/geometry/llgeo.c:30-36 */
  yl = -frs[1] * dx + frs[0] * dy;
  cc = c * c;
  ssq = sn * sn;
  sc = c * sn;
  for (sidx = 0; sidx < 2; sidx++) {
    iret[sidx] = okd_dc3d(alpha, xl, yl, xr[2], -xs[2], frs[11], -ls, ls, -ws, ws,
                          sidx == 0 ? 1.0 : 0.0, sidx == 1 ? 1.0 : 0.0, full_space, u);
    if (iret[sidx]) {
      out[sidx][0] = out[sidx][1] = out[sidx][2] = 0.0;
      continue;
    }
    /* Hooke's law (eval_okada.c:156-167) */
    iso = lambda * (u[3] + u[7] + u[11]);
    s[0][0] = iso + mu2 * u[3];
    s[0][1] = s[1][0] = mu * (u[6] + u[4]);
    s[0][2] = s[2][0] = mu * (u[9] + u[5]);
    s[1][1] = iso + mu2 * u[7];
    s[1][2] = s[2][1] = mu * (u[10] + u[8]);
    s[2][2] = iso + mu2 * u[11];
    /* rotation about z back to the global frame (rotate_mat_z, llgeo.c:79-112) */
    g[0][0] = s[0][0] * cc + s[1][1] * ssq + s[0][1] * 2.0 * sn * c;
    g[0][1] = g[1][0] = s[0][1] * (cc - ssq) + (-s[0][0] + s[1][1]) * sc;
    g[0][2] = g[2][0] = s[0][2] * c + s[1][2] * sn;
    g[1][1] = s[1][1] * cc - 2.0 * s[0][1] * sc + s[0][0] * ssq;
    g[1][2] = g[2][1] = s[1][2] * c - s[0][2] * sn;
    g[2][2] = s[2][2];
    /* traction on the receiver plane (resolve_force, geometry.c:73-82) */
    for (i = 0; i < 3; i++) {
      tr[i] = 0.0;
      for (j = 0; j < 3; j++) tr[i] += frr[2 + j] * g[i][j];
    }
    /* projections (resolve_stress_on_fault, geometry.c:84-111) */
    out[sidx][0] = tr[0] * frr[5] + tr[1] * frr[6] + tr[2] * frr[7];
    out[sidx][1] = tr[0] * frr[8] + tr[1] * frr[9] + tr[2] * frr[10];
    out[sidx][2] = tr[0] * frr[2] + tr[1] * frr[3] + tr[2] * frr[4];
  }
}

void rsfo_traction_pair(const double *xr, const double *frr, const double *xs,
                        const double *frs, double ls, double ws,
                        int full_space, double out[2][3], int iret[2])
{
  real o[2][3];
  int sidx, t;
  okd_traction_pair(xr, frr, xs, frs, ls, ws, full_space, o, iret);
  for (sidx = 0; sidx < 2; sidx++)
    for (t = 0; t < 3; t++) out[sidx][t] = OKD_TO_DOUBLE(o[sidx][t]);
}

void rsfo_traction_tensor(int n, const double *xyz, const float *strike,
                          const float *dip, const double *l, const double *w,
                          int full_space, int nthreads, double *T, int *iret)
{
  double *fr = (double *)malloc(sizeof(double) * 13 * (size_t)n);
  long long npair = (long long)n * n, pp;
  int i;
  for (i = 0; i < n; i++) rsfo_patch_frame(strike[i], dip[i], fr + 13 * i);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
  for (pp = 0; pp < npair; pp++) {
    int ir = (int)(pp / n), js = (int)(pp % n), ie[2];
    rsfo_traction_pair(xyz + 3 * ir, fr + 13 * ir, xyz + 3 * js, fr + 13 * js,
                       l[js], w[js], full_space, (double (*)[3])(T + pp * 6), ie);
    if (iret) {
      iret[2 * pp] = ie[0];
      iret[2 * pp + 1] = ie[1];
    }
  }
  free(fr);
}

int rsfo_read_patches(const char *fname, int nmax, double *xyz, float *strike,
                      float *dip, double *l, double *w, int *group)
{
  FILE *in = fopen(fname, "r");
  int n = 0;
  if (!in) return -1;
  while (n < nmax && fscanf(in, "%lf %lf %lf %f %f %lf %lf %i", xyz + 3 * n,
                            xyz + 3 * n + 1, xyz + 3 * n + 2, strike + n, dip + n,
                            l + n, w + n, group + n) == 8)
    n++;
  fclose(in);
  return n;
}
