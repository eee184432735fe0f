// Dense operator assembly: one thread per (receiver, source) pair.
//
// Replaces GenKEntries_petsc (src/interact_matrix_assembly.c:102-151) and
// the dense branch of calc_petsc_Isn_matrices
// (src/petsc_matrix_operations.c:379-503, MatScale :735).  A workgroup
// covers 256 consecutive sources for one receiver, so receiver data is
// wave-uniform and every store of a row segment is contiguous.
#include <hip/hip_runtime.h>

#include "okada_dev.h"
#include "rsfg_common.h"
#include "tri_dev.h"

namespace rsfg {

struct AsmArgs {
  const double* geo;
  int n, full, mode, G, ncomp, with_normal;
  double scale, lambda, mu, vtol;
  OkMedium med;
  long ld;
  double* K64;
  float* K32;
  bf16x2_t* Kbx2;
  bf16_t* Kb16;
};


// Unscaled tractions at receiver i from unit strike (k = 0) and dip
// (k = 1) slip on source j: T[k][0..2] = strike, dip, normal traction on
// the receiver plane.  Returns true for an edge-singular pair (T = 0).
__device__ __forceinline__ bool pair_traction(const double* __restrict__ geo,
                                              int n, int i, int j, bool full,
                                              const OkMedium& med,
                                              double lambda, double mu,
                                              double T[2][3]) {
  const double ca = geo[kGeoCa * n + j], sa = geo[kGeoSa * n + j];
  const double dx = geo[kGeoX * n + i] - geo[kGeoX * n + j];
  const double dy = geo[kGeoY * n + i] - geo[kGeoY * n + j];
  // receiver in the source's strike frame (z stays absolute)
  const double xl = ca * dx + sa * dy;
  const double yl = -sa * dx + ca * dy;
  OkGrad g;
  const bool sing =
      okada_grad(xl, yl, geo[kGeoZ * n + i], -geo[kGeoZ * n + j],
                 geo[kGeoSd * n + j], geo[kGeoCd * n + j], geo[kGeoL * n + j],
                 geo[kGeoW * n + j], full, med, g);
  if (sing) {
#pragma unroll
    for (int k = 0; k < 2; ++k) T[k][0] = T[k][1] = T[k][2] = 0.0;
    return true;
  }
  // receiver plane vectors expressed in the source frame
  double nl[3], sl[3], dl[3];
  {
    const double n0 = geo[kGeoN0 * n + i], n1 = geo[kGeoN1 * n + i];
    const double s0 = geo[kGeoS0 * n + i], s1 = geo[kGeoS1 * n + i];
    const double d0 = geo[kGeoD0 * n + i], d1 = geo[kGeoD1 * n + i];
    nl[0] = ca * n0 + sa * n1;
    nl[1] = -sa * n0 + ca * n1;
    nl[2] = geo[kGeoN2 * n + i];
    sl[0] = ca * s0 + sa * s1;
    sl[1] = -sa * s0 + ca * s1;
    sl[2] = geo[kGeoS2 * n + i];
    dl[0] = ca * d0 + sa * d1;
    dl[1] = -sa * d0 + ca * d1;
    dl[2] = geo[kGeoD2 * n + i];
  }
  const double mu2 = 2.0 * mu;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double* u = k ? g.d : g.s;
    // Hooke's law in the source frame
    const double iso = lambda * (u[0] + u[4] + u[8]);
    const double sxx = iso + mu2 * u[0];
    const double syy = iso + mu2 * u[4];
    const double szz = iso + mu2 * u[8];
    const double sxy = mu * (u[3] + u[1]);
    const double sxz = mu * (u[6] + u[2]);
    const double syz = mu * (u[7] + u[5]);
    const double t0 = sxx * nl[0] + sxy * nl[1] + sxz * nl[2];
    const double t1 = sxy * nl[0] + syy * nl[1] + syz * nl[2];
    const double t2 = sxz * nl[0] + syz * nl[1] + szz * nl[2];
    T[k][0] = t0 * sl[0] + t1 * sl[1] + t2 * sl[2];
    T[k][1] = t0 * dl[0] + t1 * dl[1] + t2 * dl[2];
    T[k][2] = t0 * nl[0] + t1 * nl[1] + t2 * nl[2];
  }
  return false;
}

// Tractions at receiver i from unit strike (k = 0) and dip (k = 1) slip on
// the triangular source j (always a half space, as eval_triangle_general
// evaluates triangles, src/eval_green.c:156-168): T[k] = (strike, dip,
// normal) components of S n on the receiver plane (resolve_force /
// resolve_stress_on_fault, src/geometry/geometry.c:73-111).  Returns true
// when the stress is not finite (T = 0, the reference's entry 0).  vtol:
// the vertical-edge tolerance of the free-surface correction.
__device__ __forceinline__ bool tri_pair_traction(const double* __restrict__ geo, int n, int i,
                                                  int j, double lambda, double mu, double vtol,
                                                  double T[2][3]) {
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = geo[(kGeoV0 + k) * n + j];
  const double P[3] = {geo[kGeoX * n + i], geo[kGeoY * n + i], geo[kGeoZ * n + i]};
  double nr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) nr[k] = geo[(kGeoN0 + k) * n + i];
  double t[2][3];
  if (!tri_unit_traction(P, v, nr, 2.0 * mu, lambda, vtol, t)) {
#pragma unroll
    for (int k = 0; k < 2; ++k) T[k][0] = T[k][1] = T[k][2] = 0.0;
    return true;
  }
  // the receiver's strike and dip vectors only now (not live during the
  // stress evaluation)
  double sr[3], dr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    sr[k] = geo[(kGeoS0 + k) * n + i];
    dr[k] = geo[(kGeoD0 + k) * n + i];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    T[k][0] = t[k][0] * sr[0] + t[k][1] * sr[1] + t[k][2] * sr[2];
    T[k][1] = t[k][0] * dr[0] + t[k][1] * dr[1] + t[k][2] * dr[2];
    T[k][2] = t[k][0] * nr[0] + t[k][1] * nr[1] + t[k][2] * nr[2];
  }
  return false;
}

// source-type dispatch by the mesh's kinds of patch (kSrcRect: rectangles
// only, kSrcTri: triangles only, kSrcMixed: both): a kernel compiles only
// the paths its mesh can reach, so its register budget is theirs
enum { kSrcRect = 0, kSrcMixed = 1, kSrcTri = 2 };
template <int SRC>
__device__ __forceinline__ bool any_pair_traction(const double* __restrict__ geo, int n, int i,
                                                  int j, bool full, const OkMedium& med,
                                                  double lambda, double mu, double vtol,
                                                  double T[2][3]) {
  if constexpr (SRC == kSrcTri) {
    return tri_pair_traction(geo, n, i, j, lambda, mu, vtol, T);
  } else {
    if (SRC == kSrcMixed && geo[kGeoTri * n + j] != 0.0)
      return tri_pair_traction(geo, n, i, j, lambda, mu, vtol, T);
    return pair_traction(geo, n, i, j, full, med, lambda, mu, T);
  }
}

static int mesh_sources(const rsfg_ctx* h) {
  return h->has_tri ? (h->has_rect ? kSrcMixed : kSrcTri) : kSrcRect;
}

__device__ __forceinline__ void store_entry(const AsmArgs& a, size_t idx,
                                            double v) {
  a.K64[idx] = v;
  if (a.K32) a.K32[idx] = static_cast<float>(v);
  if (a.Kbx2 || a.Kb16) {
    const uint32_t hi = bf16_rn(static_cast<float>(v));
    if (a.Kb16) a.Kb16[idx] = static_cast<bf16_t>(hi);
    if (a.Kbx2) {
      const double rem = v - static_cast<double>(__uint_as_float(hi << 16));
      const uint32_t lo = bf16_rn(static_cast<float>(rem));
      a.Kbx2[idx] = (hi << 16) | lo;
    }
  }
}

#ifndef RSFG_ASM_MINW
#define RSFG_ASM_MINW 1
#endif
template <int SRC>
__global__ __launch_bounds__(256, RSFG_ASM_MINW) void k_assemble(AsmArgs a) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= a.n) return;
  double T[2][3];
  any_pair_traction<SRC>(a.geo, a.n, i, j, a.full != 0, a.med, a.lambda, a.mu, a.vtol, T);
  const int n = a.n, G = a.G, nc = a.ncomp;
  const double sc = a.scale;
  switch (a.mode) {
    case RSFG_SLIP_STRIKE:
    case RSFG_SLIP_DIP: {
      const int k = (a.mode == RSFG_SLIP_DIP) ? 1 : 0;
      store_entry(a, (size_t)(i * G) * a.ld + j, T[k][k] * sc);
      if (a.with_normal) store_entry(a, (size_t)(i * G + 1) * a.ld + j, T[k][2] * -sc);
      break;
    }
    case RSFG_SLIP_RAKE: {
      const double rs0 = a.geo[kGeoR0 * n + j], rs1 = a.geo[kGeoR1 * n + j];
      const double rr0 = a.geo[kGeoR0 * n + i], rr1 = a.geo[kGeoR1 * n + i];
      const double ts = rs0 * T[0][0] + rs1 * T[1][0];
      const double tdp = rs0 * T[0][1] + rs1 * T[1][1];
      store_entry(a, (size_t)(i * G) * a.ld + j, (rr0 * ts + rr1 * tdp) * sc);
      if (a.with_normal)
        store_entry(a, (size_t)(i * G + 1) * a.ld + j, (rs0 * T[0][2] + rs1 * T[1][2]) * -sc);
      break;
    }
    default: {  // two slip components, dof 2p + c
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int k = 0; k < 2; ++k)
          store_entry(a, (size_t)(i * G + t) * a.ld + nc * j + k, T[k][t] * sc);
      if (a.with_normal) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
          store_entry(a, (size_t)(i * G + 2) * a.ld + nc * j + k, T[k][2] * -sc);
      }
      break;
    }
  }
}

// bulk traction blocks for arbitrary pairs (unscaled, internal units)
template <int SRC>
__global__ __launch_bounds__(256) void k_pairs(const double* geo, int n,
                                               long np, const int* recv,
                                               const int* src, int full,
                                               OkMedium med, double lambda,
                                               double mu, double vtol, double* out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  double T[2][3];
  any_pair_traction<SRC>(geo, n, recv[p], src[p], full != 0, med, lambda, mu, vtol, T);
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int t = 0; t < 3; ++t) out[6 * p + 3 * k + t] = T[k][t];
}

// internal elastic constants: shear modulus 1e4, Poisson 0.25
// (src/includes/properties.h:23,39; src/read_geometry.c:702-719)
static void internal_medium(OkMedium& m, double& lambda, double& mu) {
  mu = 1.0e4;
  const double nu = 0.25, mu2 = 2.0 * mu;
  lambda = (mu2 * nu) / (1.0 - 2.0 * nu);
  const double alpha = (lambda + mu) / (lambda + mu2);
  m.a1 = (1.0 - alpha) / 2.0;
  m.a2 = alpha / 2.0;
  m.a3 = (1.0 - alpha) / alpha;
  m.a4 = 1.0 - alpha;
  m.a5 = alpha;
}

int launch_assemble(rsfg_ctx* h, double* gpu_ms) {
  AsmArgs a;
  a.geo = h->d_geo;
  a.n = h->n;
  a.full = h->full_space;
  a.mode = h->slip_mode;
  a.G = h->G;
  a.ncomp = h->ncomp;
  a.with_normal = h->with_normal;
  a.scale = h->scale;
  internal_medium(a.med, a.lambda, a.mu);
  a.vtol = h->tri_vtol;
  a.ld = h->ld;
  a.K64 = h->K64;
  a.K32 = h->K32;
  a.Kbx2 = h->Kbx2;
  a.Kb16 = h->Kb16;
  dim3 grid((h->n + 255) / 256, h->n);
  RSFG_HIP(hipEventRecord(h->ev0, h->stream));
  switch (mesh_sources(h)) {
    case kSrcTri:
      hipLaunchKernelGGL(k_assemble<kSrcTri>, grid, dim3(256), 0, h->stream, a);
      break;
    case kSrcMixed:
      hipLaunchKernelGGL(k_assemble<kSrcMixed>, grid, dim3(256), 0, h->stream, a);
      break;
    default:
      hipLaunchKernelGGL(k_assemble<kSrcRect>, grid, dim3(256), 0, h->stream, a);
  }
  RSFG_HIP(hipGetLastError());
  RSFG_HIP(hipEventRecord(h->ev1, h->stream));
  RSFG_HIP(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  RSFG_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (gpu_ms) *gpu_ms = ms;
  return RSFG_OK;
}

int launch_pairs(rsfg_ctx* h, long np, const int* d_recv, const int* d_src,
                 int full, double* d_out) {
  OkMedium m;
  double lambda, mu;
  internal_medium(m, lambda, mu);
  const long nb = (np + 255) / 256;
#define RSFG_PAIRS(SRC)                                                                      \
  hipLaunchKernelGGL(k_pairs<SRC>, dim3((unsigned)nb), dim3(256), 0, h->stream, h->d_geo, h->n, \
                     np, d_recv, d_src, full, m, lambda, mu, h->tri_vtol, d_out)
  switch (mesh_sources(h)) {
    case kSrcTri: RSFG_PAIRS(kSrcTri); break;
    case kSrcMixed: RSFG_PAIRS(kSrcMixed); break;
    default: RSFG_PAIRS(kSrcRect);
  }
#undef RSFG_PAIRS
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

}  // namespace rsfg
