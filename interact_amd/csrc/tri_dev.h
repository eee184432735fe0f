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

// the reference's pi in the angle offsets and its edge tolerance
// (src/green/tdstresshs.f90 pi parameter, EPS_FOR_FORTRAN)
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

// T <- A T A^T (T symmetric, stored full)
__device__ __forceinline__ void trot(const double A[3][3], double T[3][3]) {
  double M[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = A[i][0] * T[0][j] + A[i][1] * T[1][j] + A[i][2] * T[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) T[i][j] = M[i][0] * A[j][0] + M[i][1] * A[j][1] + M[i][2] * A[j][2];
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

// Full-space strain in the global frame for unit strike slip (E[0]) and
// unit dip slip (E[1]).  Also returns the TDCS strike and dip vectors (the
// global Burgers vectors of the two unit slips).  false: point on an edge
// line in the triangle's plane.
__device__ __forceinline__ bool td_fs_unit(const double* P, const double* p1, const double* p2,
                                           const double* p3, double E[2][3][3], double* vs,
                                           double* vd) {
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
  double Et[2][3][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Et[s][i][j] = 0.0;
  for (int k = 0; k < 3; ++k) {
    // angular-dislocation frame: rotation about the normal that takes the
    // side vector to -z (get_tdcd_adcs, src/green/tddisphs.f90:826-842)
    const double c = svz[k], s = svy[k];
    const double dy = q[1] - apy[k], dz = q[2] - apz[k];
    double sA, cA;
    sincos(-kTriPi + ang[k], &sA, &cA);
    double gb[3][3][3];
    nw_adis_gradb(q[0], c * dy - s * dz, s * dy + c * dz, cA, sA, kTriNu, gb);
    // unit strike slip: b = (0, c, s) in the dislocation frame; dip: (0, -s, c)
    const double wy[2] = {c, -s}, wz[2] = {s, c};
    const double B[3][3] = {{1.0, 0.0, 0.0}, {0.0, c, s}, {0.0, -s, c}};
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      double Ea[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double gij = wy[sl] * gb[1][i][j] + wz[sl] * gb[2][i][j];
          const double gji = wy[sl] * gb[1][j][i] + wz[sl] * gb[2][j][i];
          Ea[i][j] = 0.5 * (gij + gji);
        }
      trot(B, Ea);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Et[sl][i][j] += Ea[i][j];
    }
  }
  // Burgers function of the triangle: b = e_strike (TDCS y) or e_dip (TDCS z)
  {
    double gf[3][3];
    nw_burg_grad(q[0], q[1], q[2], t1[1], t1[2], t3[1], t3[2], gf);
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      const int c = 1 + sl;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Et[sl][c][j] += 0.5 * gf[0][j];
        Et[sl][j][c] += 0.5 * gf[0][j];
      }
    }
  }
  const double Ar[3][3] = {{vn[0], vs[0], vd[0]}, {vn[1], vs[1], vd[1]}, {vn[2], vs[2], vd[2]}};
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    trot(Ar, Et[sl]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) E[sl][i][j] = Et[sl][i][j];
  }
  return true;
}

// free-surface correction on edge PA-PB for the two global Burgers vectors
// bs (strike slip) and bd (dip slip), accumulated into E (AngSetupFSC_S,
// src/green/tdstresshs.f90:371-476)
__device__ __forceinline__ void fsc_edge_unit(const double* P, const double* bs, const double* bd,
                                              const double* pa, const double* pb,
                                              double E[2][3][3]) {
  const double sv[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
  const double beta = acos(-sv[2] / sqrt(tdot(sv, sv)));
  if (fabs(beta) < kTriEps || fabs(kTriPi - beta) < kTriEps) return;
  double e1[3] = {sv[0], sv[1], 0.0};
  tunit(e1);
  // frame (e1, (-e_z) x e1, -e_z): symmetric, its own transpose
  const double A[3][3] = {{e1[0], e1[1], 0.0}, {e1[1], -e1[0], 0.0}, {0.0, 0.0, -1.0}};
  const double d[3] = {P[0] - pa[0], P[1] - pa[1], P[2] - pa[2]};
  double yA[3], yB[3], bbs[3], bbd[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    yA[i] = A[i][0] * d[0] + A[i][1] * d[1] + A[i][2] * d[2];
    yB[i] = yA[i] - (A[i][0] * sv[0] + A[i][1] * sv[1] + A[i][2] * sv[2]);
    bbs[i] = A[i][0] * bs[0] + A[i][1] * bs[1] + A[i][2] * bs[2];
    bbd[i] = A[i][0] * bd[0] + A[i][1] * bd[1] + A[i][2] * bd[2];
  }
  double sB, cB;
  sincos((beta * yA[0] >= 0.0) ? -kTriPi + beta : beta, &sB, &cB);
  double gA[3][3][3], gB[3][3][3];
  nw_fsc_gradb(yA[0], yA[1], yA[2], cB, sB, kTriNu, -pa[2], gA);
  nw_fsc_gradb(yB[0], yB[1], yB[2], cB, sB, kTriNu, -pb[2], gB);
  const double At[3][3] = {{A[0][0], A[1][0], A[2][0]}, {A[0][1], A[1][1], A[2][1]},
                           {A[0][2], A[1][2], A[2][2]}};
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const double* bb = sl ? bbd : bbs;
    double V[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double gij = 0.0, gji = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          gij += bb[c] * (gB[c][i][j] - gA[c][i][j]);
          gji += bb[c] * (gB[c][j][i] - gA[c][j][i]);
        }
        V[i][j] = 0.5 * (gij + gji);
      }
    trot(At, V);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) E[sl][i][j] += V[i][j];
  }
}

__device__ __forceinline__ void hooke_add(const double E[3][3], double mu2, double lambda,
                                          double S[3][3]) {
  const double tr = lambda * (E[0][0] + E[1][1] + E[2][2]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] += mu2 * E[i][j] + (i == j ? tr : 0.0);
}

// Stress at P for unit strike slip (S[0]) and unit dip slip (S[1]) on the
// triangle v[0..8], internal elastic constants (mu2 = 2 mu, lambda).  false
// when the stress is not finite (the reference's iret = 1, entry 0).  A
// triangle with all vertices at z = 0 gives zero stress.
__device__ __forceinline__ bool tri_unit_stress(const double* P, const double* v, double mu2,
                                                double lambda, double S[2][3][3]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) S[s][i][j] = 0.0;
  if (v[2] == 0.0 && v[5] == 0.0 && v[8] == 0.0) return true;
  double E[2][3][3], vs[3], vd[3];
  // main dislocation
  if (!td_fs_unit(P, v, v + 3, v + 6, E, vs, vd)) return false;
  double Sm[2][3][3], Sf[2][3][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Sm[s][i][j] = Sf[s][i][j] = 0.0;
  hooke_add(E[0], mu2, lambda, Sm[0]);
  hooke_add(E[1], mu2, lambda, Sm[1]);
  // harmonic correction along the three edges
  double Ef[2][3][3];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ef[s][i][j] = 0.0;
  fsc_edge_unit(P, vs, vd, v, v + 3, Ef);
  fsc_edge_unit(P, vs, vd, v + 3, v + 6, Ef);
  fsc_edge_unit(P, vs, vd, v + 6, v, Ef);
  hooke_add(Ef[0], mu2, lambda, Sf[0]);
  hooke_add(Ef[1], mu2, lambda, Sf[1]);
  // image dislocation
  const double w[9] = {v[0], v[1], -v[2], v[3], v[4], -v[5], v[6], v[7], -v[8]};
  double vsi[3], vdi[3];
  if (!td_fs_unit(P, w, w + 3, w + 6, E, vsi, vdi)) return false;
  bool ok = true;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    double Si[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    hooke_add(E[s], mu2, lambda, Si);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        S[s][i][j] = Sm[s][i][j] + Si[i][j] + Sf[s][i][j];
        ok = ok && isfinite(S[s][i][j]);
      }
  }
  return ok;
}

}  // namespace rsfg
