// C ABI of librsfg (declarations and reference mapping in include/rsfg.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gemv_dev.h"
#include "rsf_dev.h"
#include "rsfg_common.h"

namespace rsfg {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
           hipGetErrorString(e), what, file, line);
  g_err = buf;
  return RSFG_EHIP;
}

int launch_assemble(rsfg_ctx* h, double* gpu_ms);
int launch_pairs(rsfg_ctx* h, long np, const int* d_recv, const int* d_src, int full,
                 double* d_out);
int solve_device(rsfg_ctx* h, const rsfg_solver_opts* o, double t_end, rsfg_stats* st,
                 bool cont);
int stage_bench(rsfg_ctx* h, int prec, int nrep, double* ms);

constexpr double kDeg2Rad = 0.0174532925199432957692369076848861271344287188854172545609719144;
constexpr double kOkEps = 1.0e-8;
constexpr size_t kChunkMax = 512;   // largest GEMV chunk (bf16: 64 lanes x 8)

// Geometry of a triangular patch as read_geometry derives it
// (src/read_geometry.c:148-215 -> get_tri_prop_based_on_gh,
// src/green/eval_triangle_nw.c:99-140): receiver point = vertex mean
// (calc_centroid_tri, src/geometry/geometry.c:1037-1045), receiver plane
// vectors = the TDCS (normal, strike, dip) of get_tdcs_base_vectors
// (src/green/tddisphs.f90:790-817).  The rectangle fields stay zero.
static void tri_geometry(const rsfg_patch& q, int n, int i, double* g) {
  const double* v = q.vert;
  g[kGeoX * n + i] = (v[0] + v[3] + v[6]) / 3.0;
  g[kGeoY * n + i] = (v[1] + v[4] + v[7]) / 3.0;
  g[kGeoZ * n + i] = (v[2] + v[5] + v[8]) / 3.0;
  const double a[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
  const double b[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
  double vn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                  a[0] * b[1] - a[1] * b[0]};
  double len = sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
  for (int k = 0; k < 3; ++k) vn[k] /= len;
  double vs[3] = {-vn[1], vn[0], 0.0};
  if (sqrt(vs[0] * vs[0] + vs[1] * vs[1]) == 0.0) {
    vs[0] = 0.0;
    vs[1] = (v[2] > 0.0) ? -vn[2] : vn[2];
  }
  len = sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
  for (int k = 0; k < 3; ++k) vs[k] /= len;
  const double vd[3] = {vn[1] * vs[2] - vn[2] * vs[1], vn[2] * vs[0] - vn[0] * vs[2],
                        vn[0] * vs[1] - vn[1] * vs[0]};
  for (int k = 0; k < 3; ++k) {
    g[(kGeoN0 + k) * n + i] = vn[k];
    g[(kGeoS0 + k) * n + i] = vs[k];
    g[(kGeoD0 + k) * n + i] = vd[k];
  }
  g[kGeoR0 * n + i] = q.rake[0];
  g[kGeoR1 * n + i] = q.rake[1];
  g[kGeoTri * n + i] = 1.0;
  for (int k = 0; k < 9; ++k) g[(kGeoV0 + k) * n + i] = v[k];
}

int ensure_tmp(rsfg_ctx* h, size_t bytes) {
  if (bytes <= h->tmp_bytes) return RSFG_OK;
  if (h->d_tmp) hipFree(h->d_tmp);
  h->d_tmp = nullptr;
  h->tmp_bytes = 0;
  RSFG_HIP(hipMalloc(&h->d_tmp, bytes));
  h->tmp_bytes = bytes;
  return RSFG_OK;
}

const void* operator_ptr(rsfg_ctx* h, int prec) {
  if (!h->assembled) return nullptr;
  switch (prec) {
    case RSFG_PREC_FP64: return h->K64;
    case RSFG_PREC_FP32: return h->K32;
    case RSFG_PREC_BF16X2: return h->Kbx2;
    case RSFG_PREC_BF16: return (h->store & RSFG_PREC_BF16) ? h->Kb16 : nullptr;
    case RSFG_PREC_BF16NF: return h->Knf_lo || h->nf_off ? h->Kb16 : nullptr;
    default: return nullptr;
  }
}

size_t vin_elems_for(int cols) {
  return ((size_t)cols + 2 * kChunkMax + kChunkMax - 1) / kChunkMax * kChunkMax;
}

static void free_operator(rsfg_ctx* h) {
  hipFree(h->K64);
  hipFree(h->K32);
  hipFree(h->Kbx2);
  hipFree(h->Kb16);
  hipFree(h->Knf_lo);
  hipFree(h->nf_off);
  hipFree(h->nf_list);
  h->K64 = nullptr;
  h->K32 = nullptr;
  h->Kbx2 = nullptr;
  h->Kb16 = nullptr;
  h->Knf_lo = nullptr;
  h->nf_off = nullptr;
  h->nf_list = nullptr;
  h->nf_chunks = 0;
  h->assembled = false;
}

static void free_state(rsfg_ctx* h) {
  drop_graph(h);
  h->cont_ok = false;
  hipFree(h->d_state);
  hipFree(h->partials);
  hipFree(h->d_ev);
  hipFree(h->d_mon);
  hipFree(h->d_cat);
  hipFree(h->d_catr);
  hipFree(h->d_rupt);
  h->d_state = nullptr;
  h->partials = nullptr;
  h->d_ev = h->d_mon = h->d_cat = h->d_catr = nullptr;
  h->d_rupt = nullptr;
  h->ev_cap = h->mon_cap = 0;
}

int ensure_logs(rsfg_ctx* h, int ev_rows, int mon_rows) {
  // grown logs keep the rows already written (a continued solve appends)
  if (ev_rows > h->ev_cap) {
    double* p = nullptr;
    RSFG_HIP(hipMalloc(&p, (size_t)ev_rows * 8 * sizeof(double)));
    if (h->d_ev && h->ev_cap > 0)
      RSFG_HIP(hipMemcpyAsync(p, h->d_ev, (size_t)h->ev_cap * 8 * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
    RSFG_HIP(hipStreamSynchronize(h->stream));
    hipFree(h->d_ev);
    h->d_ev = p;
    // catalog rows: at most one per event row, grown with it
    double* q = nullptr;
    RSFG_HIP(hipMalloc(&q, (size_t)ev_rows * kCatCols * sizeof(double)));
    if (h->d_catr && h->ev_cap > 0)
      RSFG_HIP(hipMemcpyAsync(q, h->d_catr, (size_t)h->ev_cap * kCatCols * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
    RSFG_HIP(hipStreamSynchronize(h->stream));
    hipFree(h->d_catr);
    h->d_catr = q;
    h->ev_cap = ev_rows;
  }
  if (mon_rows > h->mon_cap) {
    double* p = nullptr;
    RSFG_HIP(hipMalloc(&p, (size_t)mon_rows * 10 * sizeof(double)));
    if (h->d_mon && h->mon_cap > 0)
      RSFG_HIP(hipMemcpyAsync(p, h->d_mon, (size_t)h->mon_cap * 10 * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
    RSFG_HIP(hipStreamSynchronize(h->stream));
    hipFree(h->d_mon);
    h->d_mon = p;
    h->mon_cap = mon_rows;
  }
  return RSFG_OK;
}

int alloc_state(rsfg_ctx* h) {
  free_state(h);
  const size_t n = h->n, dim = h->dim;
  const size_t vec = dim * n;
  h->vin_elems = vin_elems_for(h->cols);
  static_assert(PreDim<1>::v <= 12 && PreDim<2>::v <= 12, "record buffer size");
  const size_t pre = 12 * n;  // state-only RHS records (at most 12 per patch)
  const size_t total = 13 * vec + 2 * pre + 4 * h->vin_elems;
  RSFG_HIP(hipMalloc(&h->d_state, total * sizeof(double)));
  RSFG_HIP(hipMemsetAsync(h->d_state, 0, total * sizeof(double), h->stream));
  double* p = h->d_state;
  for (int i = 0; i < 2; ++i) { h->Xb[i] = p; p += vec; }
  for (int i = 0; i < 2; ++i) { h->Yb[i] = p; p += vec; }
  for (int i = 0; i < 8; ++i) { h->Stg[i] = p; p += vec; }
  h->Xmon = p; p += vec;
  for (int i = 0; i < 2; ++i) { h->Pb[i] = p; p += pre; }
  for (int i = 0; i < 2; ++i) { h->vin[i] = p; p += h->vin_elems; }
  for (int i = 0; i < 2; ++i) { h->vxb[i] = p; p += h->vin_elems; }
  h->partial_count = (size_t)(n + 64) * kNRed;   // per-patch or per-wave columns
  RSFG_HIP(hipMalloc(&h->partials, h->partial_count * sizeof(double)));
  {
    const int rc = ensure_logs(h, kMaxEv, kMaxMon);
    if (rc != RSFG_OK) return rc;
  }
  RSFG_HIP(hipMalloc(&h->d_cat, 2 * n * sizeof(double)));
  RSFG_HIP(hipMalloc(&h->d_rupt, n * sizeof(int)));
  h->xcur = 0;
  return RSFG_OK;
}

// ---------------------------------------------------------------- kernels
template <typename VT>
__global__ void k_to_vt(const double* __restrict__ x, VT* __restrict__ y, int n, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npad) y[i] = (i < n) ? static_cast<VT>(x[i]) : VT(0);
}

// R rows per wave, four waves per workgroup
template <typename TK, int R>
__global__ __launch_bounds__(256) void k_matvec(const TK* __restrict__ K, long ld, int rows,
                                                int cols,
                                                const typename KTraits<TK>::VT* __restrict__ x,
                                                double* __restrict__ y) {
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
  double s[R];
  gemv_wave<TK, R>(K, ld, row0, rows, cols, x, s);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row0 + r < rows) y[row0 + r] = s[r];
  }
}

// bf16 + near-field lo product: one patch (its G rows) per wave
template <int G>
__global__ __launch_bounds__(256) void k_matvec_nf(const bf16_t* __restrict__ Khi,
                                                   const bf16_t* __restrict__ lo,
                                                   const int* __restrict__ off,
                                                   const uint16_t* __restrict__ list, long ld,
                                                   int n, int cols, const float* __restrict__ x,
                                                   double* __restrict__ y) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int pc = min(p, n - 1);
  const int o = off[pc];
  double s[G];
  gemv_nf_wave<G>(Khi, lo + (size_t)o * G * kNfChunk, list + o, off[pc + 1] - o, ld, pc * G,
                  n * G, cols, x, s);
  if ((threadIdx.x & 63) == 0 && p < n) {
#pragma unroll
    for (int r = 0; r < G; ++r) y[p * G + r] = s[r];
  }
}

// energy (sum of squares) of the G rows of patch p in column chunk c of
// 512 entries: one wave per (p, c)
__global__ __launch_bounds__(256) void k_chunk_energy(const double* __restrict__ K, long ld,
                                                      int n, int G, int cols, int nch,
                                                      double* __restrict__ E) {
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= (long)n * nch) return;
  const int p = (int)(w / nch), c = (int)(w % nch);
  const int l8 = (threadIdx.x & 63) * 8;
  double s = 0.0;
  for (int r = 0; r < G; ++r) {
    const double* row = K + (size_t)(p * G + r) * ld;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int col = c * kNfChunk + l8 + e;
      const double v = (col < cols) ? row[col] : 0.0;
      s = fma(v, v, s);
    }
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) E[w] = s;
}

// lo blocks of the listed chunks: lo = RN-bf16 of (K - hi), hi = RN-bf16 of
// the FP32 value (the pair stored by the bf16 hi+lo format); one wave per
// listed chunk q (owner[q] = its patch)
__global__ __launch_bounds__(256) void k_fill_nf_lo(const double* __restrict__ K, long ld, int G,
                                                    int cols, long nq,
                                                    const int* __restrict__ owner,
                                                    const uint16_t* __restrict__ list,
                                                    bf16_t* __restrict__ lo) {
  const long q = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  const int p = owner[q], c = list[q];
  const int l8 = (threadIdx.x & 63) * 8;
  for (int r = 0; r < G; ++r) {
    const double* row = K + (size_t)(p * G + r) * ld;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int col = c * kNfChunk + l8 + e;
      const double v = (col < cols) ? row[col] : 0.0;
      const uint32_t hi = bf16_rn(static_cast<float>(v));
      const double rem = v - static_cast<double>(__uint_as_float(hi << 16));
      lo[((size_t)q * G + r) * kNfChunk + l8 + e] = static_cast<bf16_t>(bf16_rn(static_cast<float>(rem)));
    }
  }
}

// y_part[split][j] = sum over rows of this split of K[i][j] x[i]
template <typename TK>
__global__ __launch_bounds__(256) void k_matvec_t(const TK* __restrict__ K, long ld, int rows,
                                                  int cols, const double* __restrict__ x,
                                                  int rows_per_split, double* __restrict__ part) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = blockIdx.y * rows_per_split;
  const int i1 = min(rows, i0 + rows_per_split);
  if (j >= cols) return;
  double acc = 0.0;
  for (int i = i0; i < i1; ++i) {
    double kv;
    const TK e = K[(size_t)i * ld + j];
    if constexpr (sizeof(TK) == 8) kv = static_cast<double>(e);
    else if constexpr (sizeof(TK) == 4 && std::is_same<TK, float>::value) kv = e;
    else if constexpr (sizeof(TK) == 4) kv = (double)bf_hi(e) + (double)bf_lo(e);
    else kv = (double)__uint_as_float(((uint32_t)e) << 16);
    acc = fma(kv, x[i], acc);
  }
  part[(size_t)blockIdx.y * cols + j] = acc;
}

__global__ void k_sum_splits(const double* __restrict__ part, int nsplit, int cols,
                             double* __restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * cols + j];
  y[j] = s;
}

__global__ void k_pick_shear(const double* __restrict__ yall, int n, int nc, int G,
                             double* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n * nc) return;
  y[r] = yall[(size_t)(r / nc) * G + (r % nc)];
}

__global__ void k_gather(const double* __restrict__ K, long ld, int m, const int* rows, int k,
                         const int* cols, double* out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)m * k) return;
  const int r = (int)(t % m), c = (int)(t / m);
  out[t] = K[(size_t)rows[r] * ld + cols[c]];
}

// velocities of a SoA state (for rsfg_rhs)
template <typename VT, int NC>
__global__ void k_vel(const double* __restrict__ Y, const double* __restrict__ coef,
                      FricParams fp, int n, VT* __restrict__ v, int* __restrict__ ok) {
  constexpr int D = 2 * NC + 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double y[D], vv[2];
  for (int f = 0; f < D; ++f) y[f] = Y[f * n + p];
  PatchCoef c{coef[p], coef[n + p], coef[2 * n + p]};
  rsf_patch_vel<NC>(fp, c.a, y, vv);
  for (int k = 0; k < NC; ++k) v[NC * p + k] = static_cast<VT>(vv[k]);
  ok[p] = rsf_patch_ok<NC>(fp, c, y) ? 1 : 0;
}

template <typename TK, int NC, int G>
__global__ __launch_bounds__(256) void k_rhs(const TK* __restrict__ K, long ld, int rows, int cols,
                                             const typename KTraits<TK>::VT* __restrict__ v,
                                             const double* __restrict__ Y,
                                             const double* __restrict__ coef,
                                             const double* __restrict__ sinc, FricParams fp, int n,
                                             double* __restrict__ F) {
  // one patch (its G rows) per wave, as in the fused stage kernel
  constexpr int D = 2 * NC + 2;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  double rowsum[G];
  gemv_wave<TK, G>(K, ld, min(p, n - 1) * G, rows, cols, v, rowsum);
  if ((threadIdx.x & 63) != 0 || p >= n) return;
  double y[D], f[D], td[NC], sn[NC];
  for (int k = 0; k < D; ++k) y[k] = Y[k * n + p];
  for (int k = 0; k < NC; ++k) {
    td[k] = rowsum[k];
    sn[k] = sinc[NC * p + k];
  }
  const double sdn = (G > NC) ? rowsum[NC] : 0.0;
  PatchCoef c{coef[p], coef[n + p], coef[2 * n + p]};
  rsf_patch_rhs<NC>(fp, c, y, td, sdn, sn, sinc[NC * n + p], f);
  for (int k = 0; k < D; ++k) F[k * n + p] = f[k];
}

template <typename TK>
static int matvec_t_launch(rsfg_ctx* h, const TK* K, const double* d_x, double* d_y,
                           double* d_part, int nsplit) {
  const int rows = h->rows_all, cols = h->cols;
  const int rps = (rows + nsplit - 1) / nsplit;
  dim3 grid((cols + 255) / 256, nsplit);
  hipLaunchKernelGGL(k_matvec_t<TK>, grid, dim3(256), 0, h->stream, K, (long)h->ld, rows, cols,
                     d_x, rps, d_part);
  hipLaunchKernelGGL(k_sum_splits, dim3((cols + 255) / 256), dim3(256), 0, h->stream, d_part,
                     nsplit, cols, d_y);
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

// Persistent product of the FP32-input formats: one workgroup of kPW waves
// per CU copies x into LDS once, then each wave streams upw consecutive
// units against that copy (a unit is R = 2 rows of the plain formats, or
// one patch's G rows and its near-field lo chunks for bf16nf).  Every row
// sum is formed in the order of k_matvec / k_matvec_nf (same bits).
template <typename TK, int R, bool CB>
__global__ __launch_bounds__(64 * kPW, 1) void k_matvec_p(const void* __restrict__ Kv,
                                                         const bf16_t* __restrict__ lo,
                                                         const int* __restrict__ off,
                                                         const uint16_t* __restrict__ list,
                                                         long ld, int rows, int cols,
                                                         const float* __restrict__ x,
                                                         double* __restrict__ y, int upw) {
  using TH = typename std::conditional<std::is_same<TK, bf16nf_t>::value, bf16_t, TK>::type;
  __shared__ __attribute__((aligned(16))) float xs[kXL];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nunit = (rows + R - 1) / R;
  const int ub = min((blockIdx.x * kPW + wave) * upw, nunit), ue = min(ub + upw, nunit);
  const int nch = lds_x_floats<TH>(cols) / chunk_elems<TH>();
  // the wave's rows over the chunks [c0, c1) of x, copied into LDS; a row's
  // block sums are added in block order by the lane that writes it
  auto block = [&](int c0, int c1) {
    x_to_lds<TH>(x, c0, c1, xs);
    __syncthreads();
    for (int u = ub; u < ue; ++u) {
      double acc[R];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = 0.0;
      gemv_lds_acc<TH, R, 2>(static_cast<const TH*>(Kv), ld, u * R, rows, c0, c1, xs, acc);
      if constexpr (std::is_same<TK, bf16nf_t>::value)
        gemv_nf_lo_lds<R>(lo + (size_t)off[u] * R * kNfChunk, list + off[u],
                          off[u + 1] - off[u], c0, CB ? c1 : INT_MAX, xs, acc);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double tot = wave_sum(acc[r]);
        if (lane == 0 && u * R + r < rows) y[u * R + r] = (c0 == 0) ? tot : y[u * R + r] + tot;
      }
    }
  };
  if constexpr (CB) {
    // x longer than the LDS copy: column blocks of lds_block_chunks chunks
    constexpr int BC = lds_block_chunks<TH>();
    for (int c0 = 0; c0 < nch; c0 += BC) {
      if (c0 > 0) __syncthreads();   // every wave is done with the previous block
      block(c0, min(c0 + BC, nch));
    }
  } else {
    block(0, nch);
  }
}

// Launches k_matvec_p when the format takes FP32 x and the units spread
// over the CUs; returns false (nothing launched) otherwise.
template <typename TK, int R>
static bool matvec_persistent(rsfg_ctx* h, const void* K, const float* x, double* y) {
  const char* e = getenv("RSFG_STAGE_PERSIST");
  if (e && e[0] == '0') return false;
  int nwg = 0, upw = 0;
  if (!persist_split(h->ncu, (h->rows_all + R - 1) / R, 64, &nwg, &upw)) return false;
  using TH = typename std::conditional<std::is_same<TK, bf16nf_t>::value, bf16_t, TK>::type;
  if (lds_blocks<TH>(h->cols) > 1)
    hipLaunchKernelGGL((k_matvec_p<TK, R, true>), dim3(nwg), dim3(64 * kPW), 0, h->stream, K,
                       h->Knf_lo, h->nf_off, h->nf_list, (long)h->ld, h->rows_all, h->cols, x, y,
                       upw);
  else
    hipLaunchKernelGGL((k_matvec_p<TK, R, false>), dim3(nwg), dim3(64 * kPW), 0, h->stream, K,
                       h->Knf_lo, h->nf_off, h->nf_list, (long)h->ld, h->rows_all, h->cols, x, y,
                       upw);
  return true;
}

template <typename TK>
static int matvec_launch(rsfg_ctx* h, const TK* K, const void* d_xvt, double* d_y) {
  constexpr int R = 2;
  using VT = typename KTraits<TK>::VT;
  if constexpr (std::is_same<VT, float>::value) {
    if (matvec_persistent<TK, R>(h, K, static_cast<const float*>(d_xvt), d_y)) {
      RSFG_HIP(hipGetLastError());
      return RSFG_OK;
    }
  }
  const int nb = (h->rows_all + 4 * R - 1) / (4 * R);
  hipLaunchKernelGGL((k_matvec<TK, R>), dim3(nb), dim3(256), 0, h->stream, K, (long)h->ld,
                     h->rows_all, h->cols, static_cast<const VT*>(d_xvt), d_y);
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

static int matvec_any(rsfg_ctx* h, int prec, const void* d_xvt, double* d_y) {
  switch (prec) {
    case RSFG_PREC_FP64: return matvec_launch<double>(h, h->K64, d_xvt, d_y);
    case RSFG_PREC_FP32: return matvec_launch<float>(h, h->K32, d_xvt, d_y);
    case RSFG_PREC_BF16X2: return matvec_launch<bf16x2_t>(h, h->Kbx2, d_xvt, d_y);
    case RSFG_PREC_BF16: return matvec_launch<bf16_t>(h, h->Kb16, d_xvt, d_y);
    case RSFG_PREC_BF16NF: {
      const dim3 grid((h->n + 3) / 4), block(256);
      const float* x = static_cast<const float*>(d_xvt);
      bool done = false;
      if (h->G == 1) done = matvec_persistent<bf16nf_t, 1>(h, h->Kb16, x, d_y);
      else if (h->G == 2) done = matvec_persistent<bf16nf_t, 2>(h, h->Kb16, x, d_y);
      else done = matvec_persistent<bf16nf_t, 3>(h, h->Kb16, x, d_y);
      if (done) {
        RSFG_HIP(hipGetLastError());
        return RSFG_OK;
      }
#define RSFG_NF(GG)                                                                            \
  hipLaunchKernelGGL(k_matvec_nf<GG>, grid, block, 0, h->stream, h->Kb16, h->Knf_lo, h->nf_off, \
                     h->nf_list, (long)h->ld, h->n, h->cols, x, d_y)
      if (h->G == 1) RSFG_NF(1);
      else if (h->G == 2) RSFG_NF(2);
      else RSFG_NF(3);
#undef RSFG_NF
      RSFG_HIP(hipGetLastError());
      return RSFG_OK;
    }
  }
  set_error("matvec: bad precision");
  return RSFG_EINVAL;
}

// x (FP64, device) -> the operator's input format in vin[slot], zero padded
static int stage_input(rsfg_ctx* h, int prec, const double* d_x, void* dst) {
  const int npad = (int)h->vin_elems;
  if (prec == RSFG_PREC_FP64)
    hipLaunchKernelGGL(k_to_vt<double>, dim3((npad + 255) / 256), dim3(256), 0, h->stream, d_x,
                       static_cast<double*>(dst), h->cols, npad);
  else
    hipLaunchKernelGGL(k_to_vt<float>, dim3((npad + 255) / 256), dim3(256), 0, h->stream, d_x,
                       static_cast<float*>(dst), h->cols, npad);
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

static int prec_ok(rsfg_ctx* h, int prec) {
  return operator_ptr(h, prec) != nullptr;
}

// Staging buffer of rsfg_matvec_dev_noalloc: the padded product input and,
// with normal rows, the product over all rows.  One copy per GPU, held until
// the process ends; every call on a GPU shares it.  In host code the name
// `scratch` is only a stand-in for the symbol: the device address of each
// GPU's copy comes from hipGetSymbolAddress (scratch_ptr below) and only that
// address is passed to kernels.
constexpr size_t kScratchDoubles = 65536;   // 512 KiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices] = {nullptr};

// the device address of `scratch` on `device`, looked up once per device
// with that device current (the caller has made it current)
static int scratch_ptr(int device, double** out) {
  RSFG_CHECK(device >= 0 && device < kMaxDevices, RSFG_EINVAL, "scratch: device index");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return RSFG_OK;
}

// Near-field chunk lists of the bf16 + lo format from the assembled FP64
// operator: per patch, chunks in decreasing energy until the chunks left
// out hold at most tol^2 of the patch rows' energy; then the lo blocks.
static int build_nf(rsfg_ctx* h, double tol) {
  const int n = h->n, G = h->G, cols = h->cols;
  const int nch = (cols + kNfChunk - 1) / kNfChunk;
  RSFG_CHECK(nch <= 65536, RSFG_EINVAL, "rsfg_assemble: too many columns for the nf format");
  const size_t ne = (size_t)n * nch;
  int rc = ensure_tmp(h, ne * sizeof(double));
  if (rc) return rc;
  const long waves = (long)ne;
  hipLaunchKernelGGL(k_chunk_energy, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, h->stream,
                     h->K64, (long)h->ld, n, G, cols, nch, h->d_tmp);
  RSFG_HIP(hipGetLastError());
  std::vector<double> E(ne);
  RSFG_HIP(hipMemcpyAsync(E.data(), h->d_tmp, ne * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  std::vector<int> off(n + 1, 0), owner;
  std::vector<uint16_t> list;
  std::vector<int> order(nch);
  for (int p = 0; p < n; ++p) {
    const double* e = E.data() + (size_t)p * nch;
    double total = 0.0;
    for (int c = 0; c < nch; ++c) total += e[c];
    for (int c = 0; c < nch; ++c) order[c] = c;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return e[a] > e[b]; });
    double left = total;
    std::vector<uint16_t> mine;
    for (int k = 0; k < nch && left > tol * tol * total; ++k) {
      mine.push_back((uint16_t)order[k]);
      left -= e[order[k]];
    }
    std::sort(mine.begin(), mine.end());
    for (uint16_t c : mine) {
      list.push_back(c);
      owner.push_back(p);
    }
    off[p + 1] = (int)list.size();
  }
  h->nf_chunks = (long)list.size();
  RSFG_HIP(hipMalloc(&h->nf_off, sizeof(int) * (n + 1)));
  RSFG_HIP(hipMemcpy(h->nf_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  const size_t nq = std::max<size_t>(1, list.size());
  RSFG_HIP(hipMalloc(&h->nf_list, sizeof(uint16_t) * nq));
  RSFG_HIP(hipMalloc(&h->Knf_lo, sizeof(bf16_t) * nq * G * kNfChunk));
  if (!list.empty()) {
    RSFG_HIP(hipMemcpy(h->nf_list, list.data(), sizeof(uint16_t) * list.size(),
                       hipMemcpyHostToDevice));
    rc = ensure_tmp(h, sizeof(int) * owner.size());
    if (rc) return rc;
    RSFG_HIP(hipMemcpy(h->d_tmp, owner.data(), sizeof(int) * owner.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fill_nf_lo, dim3((unsigned)((list.size() + 3) / 4)), dim3(256), 0,
                       h->stream, h->K64, (long)h->ld, G, cols, (long)list.size(),
                       reinterpret_cast<const int*>(h->d_tmp), h->nf_list, h->Knf_lo);
    RSFG_HIP(hipGetLastError());
    RSFG_HIP(hipStreamSynchronize(h->stream));
  }
  return RSFG_OK;
}

}  // namespace rsfg

using namespace rsfg;

extern "C" {

const char* rsfg_last_error(void) { return g_err.c_str(); }

const char* rsfg_version(void) { return "librsfg 0.1 (gfx950)"; }

int rsfg_device_count(int* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (count) *count = c;
  return RSFG_OK;
}

int rsfg_read_patch_file(const char* fname, rsfg_patch* patches, int* n) {
  RSFG_CHECK(fname && n && *n >= 0, RSFG_EINVAL, "rsfg_read_patch_file: bad arguments");
  FILE* in = fopen(fname, "r");
  RSFG_CHECK(in != nullptr, RSFG_EINVAL, std::string("cannot open ") + fname);
  const int cap = patches ? *n : 0;
  int k = 0;
  rsfg_patch p;
  memset(&p, 0, sizeof p);
  int bad = 0;
  while (fscanf(in, "%lf %lf %lf %f %f %lf %lf %i", &p.x, &p.y, &p.z, &p.strike, &p.dip, &p.l,
                &p.w, &p.group) == 8) {
    p.rake[0] = 1.0;
    p.rake[1] = 0.0;
    for (int j = 0; j < 9; ++j) p.vert[j] = 0.0;
    if (p.l < 0 && p.w < 0) {
      // triangle: nine vertex coordinates follow (src/read_geometry.c:148-175)
      if (fscanf(in, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", p.vert, p.vert + 1, p.vert + 2,
                 p.vert + 3, p.vert + 4, p.vert + 5, p.vert + 6, p.vert + 7, p.vert + 8) != 9) {
        bad = 1;
        break;
      }
    } else if (p.l <= 0 || p.w <= 0) {
      bad = 2;   // point sources, 2-D segments and irregular quads are not supported
      break;
    }
    if (k < cap) patches[k] = p;
    if (k == INT_MAX) {
      bad = 3;
      break;
    }
    ++k;
  }
  // as in the reference, reading ends at the first line that does not parse
  // as a patch
  fclose(in);
  *n = k;
  RSFG_CHECK(bad != 1, RSFG_EINVAL,
             "rsfg_read_patch_file: truncated triangle line (patch " + std::to_string(k) + ")");
  RSFG_CHECK(bad != 2, RSFG_EINVAL,
             "rsfg_read_patch_file: only rectangles and triangles are supported (patch " +
                 std::to_string(k) + ")");
  RSFG_CHECK(bad != 3, RSFG_EINVAL, "rsfg_read_patch_file: too many patches");
  RSFG_CHECK(!patches || k <= cap, RSFG_EINVAL,
             std::string("rsfg_read_patch_file: ") + fname + " holds " + std::to_string(k) +
                 " patches, more than the capacity " + std::to_string(cap));
  return RSFG_OK;
}

int rsfg_create(int device, const rsfg_patch* pt, int n, rsfg_handle** out) {
  RSFG_CHECK(out && pt && n > 0, RSFG_EINVAL, "rsfg_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  RSFG_HIP(hipGetDeviceCount(&ndev));
  RSFG_CHECK(device >= 0 && device < ndev, RSFG_EINVAL, "rsfg_create: no such device");
  for (int i = 0; i < n; ++i) {
    const bool tri = pt[i].l < 0 && pt[i].w < 0;
    RSFG_CHECK(tri || (pt[i].l > 0 && pt[i].w > 0), RSFG_EINVAL,
               "rsfg_create: rectangles need positive half length and width "
               "(both negative marks a triangle)");
    if (tri) {
      RSFG_CHECK(pt[i].vert[2] <= 0.0 && pt[i].vert[5] <= 0.0 && pt[i].vert[8] <= 0.0,
                 RSFG_EINVAL, "rsfg_create: triangle vertex above the surface (z > 0)");
    }
  }
  RSFG_HIP(hipSetDevice(device));
  {
    // the static staging array of rsfg_matvec_dev_noalloc on this device
    double* sp = nullptr;
    const int rc = scratch_ptr(device, &sp);
    if (rc != RSFG_OK) return rc;
  }
  rsfg_ctx* h = new rsfg_ctx();
  h->device = device;
  if (hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    h->ncu = 0;   // the persistent kernels stay off
  h->n = n;
  h->h_geo = new double[(size_t)kGeoFields * n];
  double* g = h->h_geo;
  for (int i = 0; i < (int)kGeoFields * n; ++i) g[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    const rsfg_patch& q = pt[i];
    if (q.l < 0 && q.w < 0) {
      tri_geometry(q, n, i, g);
      h->has_tri = true;
      continue;
    }
    h->has_rect = true;
    // angle range fix-up of check_fault_angles (src/geometry/geometry.c:985-1015)
    double dip = (double)q.dip, strike = (double)q.strike;
    if (dip > 90.0) { dip = 180.0 - dip; strike += 180.0; }
    if (dip < 0.0) { dip = -dip; strike += 180.0; }
    if (strike >= 360.0) strike -= 360.0;
    if (strike < 0.0) strike = 360.0 + strike;
    const float fdip = (float)dip, fstrike = (float)strike;
    double sa, ca, sdp, cdp;
    sincos((90.0 - (double)fstrike) * kDeg2Rad, &sa, &ca);
    sincos((double)fdip * kDeg2Rad, &sdp, &cdp);
    g[kGeoX * n + i] = q.x;
    g[kGeoY * n + i] = q.y;
    g[kGeoZ * n + i] = q.z;
    g[kGeoCa * n + i] = ca;
    g[kGeoSa * n + i] = sa;
    // DC3D's own dip constants (DCCON0, src/green/dc3d.F:1161-1196)
    double sd = sdp, cd = cdp;
    if (fabs(cd) < kOkEps) {
      cd = 0.0;
      if (sd > 0.0) sd = 1.0;
      if (sd < 0.0) sd = -1.0;
    }
    g[kGeoSd * n + i] = sd;
    g[kGeoCd * n + i] = cd;
    g[kGeoL * n + i] = q.l;
    g[kGeoW * n + i] = q.w;
    // calc_quad_base_vecs (src/geometry/geometry.c:128-148)
    g[kGeoS0 * n + i] = ca;
    g[kGeoS1 * n + i] = sa;
    g[kGeoS2 * n + i] = 0.0;
    g[kGeoN0 * n + i] = sa * sdp;
    g[kGeoN1 * n + i] = -ca * sdp;
    g[kGeoN2 * n + i] = cdp;
    g[kGeoD0 * n + i] = -sa * cdp;
    g[kGeoD1 * n + i] = ca * cdp;
    g[kGeoD2 * n + i] = sdp;
    g[kGeoR0 * n + i] = q.rake[0];
    g[kGeoR1 * n + i] = q.rake[1];
  }
  // catalog area per patch, 4 l w as the reference sums it
  // (src/rsf_output.c:1117); a triangle carries l = w = sqrt(its area)
  // (src/green/eval_triangle_nw.c:120-121), so it counts 4 times its area
  std::vector<double> area(n);
  for (int i = 0; i < n; ++i) {
    const rsfg_patch& q = pt[i];
    if (q.l < 0 && q.w < 0) {
      const double* v = q.vert;
      const double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
      const double e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
      const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                   cz = e1[0] * e2[1] - e1[1] * e2[0];
      area[i] = 4.0 * 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    } else {
      area[i] = 4.0 * q.l * q.w;
    }
  }
  hipError_t e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipMalloc(&h->d_area, sizeof(double) * n);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_area, area.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&h->d_geo, sizeof(double) * kGeoFields * n);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_geo, g, sizeof(double) * kGeoFields * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_ctl, sizeof(DevCtl), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&h->d_ctl, sizeof(DevCtl));
  if (e != hipSuccess) {
    int rc = hip_fail(e, "rsfg_create", __FILE__, __LINE__);
    rsfg_destroy(h);
    return rc;
  }
  *out = h;
  return RSFG_OK;
}

int rsfg_create_member(rsfg_handle* geom, rsfg_handle** out) {
  RSFG_CHECK(geom && out, RSFG_EINVAL, "rsfg_create_member: null argument");
  RSFG_CHECK(!geom->borrowed, RSFG_EINVAL, "rsfg_create_member: geom is itself a member");
  RSFG_CHECK(geom->assembled, RSFG_ESTATE, "rsfg_create_member: assemble the geometry first");
  *out = nullptr;
  RSFG_HIP(hipSetDevice(geom->device));
  rsfg_ctx* h = new rsfg_ctx();
  h->device = geom->device;
  h->ncu = geom->ncu;
  h->n = geom->n;
  h->borrowed = true;
  h->d_geo = geom->d_geo;
  h->d_area = geom->d_area;
  h->assembled = true;
  h->has_tri = geom->has_tri;
  h->has_rect = geom->has_rect;
  h->slip_mode = geom->slip_mode;
  h->ncomp = geom->ncomp;
  h->dim = geom->dim;
  h->with_normal = geom->with_normal;
  h->full_space = geom->full_space;
  h->store = geom->store;
  h->G = geom->G;
  h->rows_all = geom->rows_all;
  h->cols = geom->cols;
  h->ld = geom->ld;
  h->scale = geom->scale;
  h->K64 = geom->K64;
  h->K32 = geom->K32;
  h->Kbx2 = geom->Kbx2;
  h->Kb16 = geom->Kb16;
  h->kslack = geom->kslack;
  h->Knf_lo = geom->Knf_lo;
  h->nf_off = geom->nf_off;
  h->nf_list = geom->nf_list;
  h->nf_chunks = geom->nf_chunks;
  h->nf_tol = geom->nf_tol;
  hipError_t e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_ctl, sizeof(DevCtl), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&h->d_ctl, sizeof(DevCtl));
  if (e != hipSuccess) {
    int rc = hip_fail(e, "rsfg_create_member", __FILE__, __LINE__);
    rsfg_destroy(h);
    return rc;
  }
  *out = h;
  return RSFG_OK;
}

int rsfg_destroy(rsfg_handle* h) {
  if (!h) return RSFG_OK;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->borrowed) {
    // the geometry handle owns the patches and the operator
    h->K64 = nullptr;
    h->K32 = nullptr;
    h->Kbx2 = nullptr;
    h->Kb16 = nullptr;
    h->Knf_lo = nullptr;
    h->nf_off = nullptr;
    h->nf_list = nullptr;
    h->d_geo = nullptr;
    h->d_area = nullptr;
  }
  free_operator(h);
  free_state(h);
  hipFree(h->d_geo);
  hipFree(h->d_area);
  hipFree(h->d_coef);
  hipFree(h->d_sinc);
  hipFree(h->d_ctl);
  hipFree(h->d_tmp);
  hipFree(h->d_hlog);
  drop_graph(h);
  if (h->h_ctl) hipHostFree(h->h_ctl);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete[] h->h_geo;
  delete h;
  return RSFG_OK;
}

int rsfg_assemble(rsfg_handle* h, const rsfg_assembly* a, double* gpu_ms) {
  RSFG_CHECK(h && a, RSFG_EINVAL, "rsfg_assemble: null argument");
  const int m = a->slip_mode;
  RSFG_CHECK(m == RSFG_SLIP_STRIKE || m == RSFG_SLIP_DIP || m == RSFG_SLIP_RAKE ||
                 m == RSFG_SLIP_VECTOR,
             RSFG_EINVAL, "rsfg_assemble: unknown slip mode");
  RSFG_CHECK(!h->borrowed, RSFG_ESTATE,
             "rsfg_assemble: a member handle shares its geometry's operator");
  RSFG_HIP(hipSetDevice(h->device));
  free_operator(h);
  h->slip_mode = m;
  h->ncomp = (m == RSFG_SLIP_VECTOR) ? 2 : 1;
  h->with_normal = a->with_normal ? 1 : 0;
  h->full_space = a->full_space ? 1 : 0;
  h->store = a->store | RSFG_PREC_FP64;
  h->scale = a->scale;
  h->G = h->ncomp + h->with_normal;
  h->rows_all = h->G * h->n;
  h->cols = h->ncomp * h->n;
  h->ld = (h->cols + 63) / 64 * 64;
  h->kslack = 2 * kChunkMax;
  const size_t ne = (size_t)h->rows_all * h->ld + h->kslack;
  RSFG_HIP(hipMalloc(&h->K64, ne * sizeof(double)));
  RSFG_HIP(hipMemsetAsync(h->K64, 0, ne * sizeof(double), h->stream));
  if (h->store & RSFG_PREC_FP32) {
    RSFG_HIP(hipMalloc(&h->K32, ne * sizeof(float)));
    RSFG_HIP(hipMemsetAsync(h->K32, 0, ne * sizeof(float), h->stream));
  }
  if (h->store & RSFG_PREC_BF16X2) {
    RSFG_HIP(hipMalloc(&h->Kbx2, ne * sizeof(bf16x2_t)));
    RSFG_HIP(hipMemsetAsync(h->Kbx2, 0, ne * sizeof(bf16x2_t), h->stream));
  }
  if (h->store & (RSFG_PREC_BF16 | RSFG_PREC_BF16NF)) {   // the nf format streams the hi part
    RSFG_HIP(hipMalloc(&h->Kb16, ne * sizeof(bf16_t)));
    RSFG_HIP(hipMemsetAsync(h->Kb16, 0, ne * sizeof(bf16_t), h->stream));
  }
  int rc = launch_assemble(h, gpu_ms);
  if (rc != RSFG_OK) return rc;
  if (h->store & RSFG_PREC_BF16NF) {
    h->nf_tol = a->nf_tol > 0.0 ? a->nf_tol : 1.0 / 64.0;
    rc = build_nf(h, h->nf_tol);
    if (rc != RSFG_OK) return rc;
  }
  h->assembled = true;
  h->have_loading = false;
  h->have_state = false;
  h->dim = 2 * h->ncomp + 2;
  return RSFG_OK;
}

int rsfg_operator_dims(rsfg_handle* h, int* rows, int* cols, int* ld) {
  RSFG_CHECK(h && h->assembled, RSFG_ESTATE, "rsfg_operator_dims: not assembled");
  if (rows) *rows = h->ncomp * h->n;
  if (cols) *cols = h->cols;
  if (ld) *ld = h->ld;
  return RSFG_OK;
}

int rsfg_operator_bytes(rsfg_handle* h, int prec, double* bytes) {
  RSFG_CHECK(h && bytes, RSFG_EINVAL, "rsfg_operator_bytes: null argument");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_operator_bytes: precision not assembled");
  const double ent = (double)h->rows_all * h->cols;
  switch (prec) {
    case RSFG_PREC_FP64: *bytes = 8.0 * ent; break;
    case RSFG_PREC_FP32:
    case RSFG_PREC_BF16X2: *bytes = 4.0 * ent; break;
    case RSFG_PREC_BF16: *bytes = 2.0 * ent; break;
    default: *bytes = 2.0 * ent + 2.0 * (double)h->nf_chunks * h->G * kNfChunk; break;
  }
  return RSFG_OK;
}

// shear row r of the operator -> row of the combined (shear + normal) storage
static int shear_row(rsfg_ctx* h, int r) { return (r / h->ncomp) * h->G + (r % h->ncomp); }

int rsfg_entries(rsfg_handle* h, int m, const int* rows, int k, const int* cols, double* out) {
  RSFG_CHECK(h && h->assembled, RSFG_ESTATE, "rsfg_entries: not assembled");
  RSFG_CHECK(m >= 0 && k >= 0 && rows && cols && out, RSFG_EINVAL, "rsfg_entries: bad args");
  if (m == 0 || k == 0) return RSFG_OK;
  const int nshear = h->ncomp * h->n;
  const int nrows = nshear + (h->with_normal ? h->n : 0);
  std::vector<int> rr(m), cc(k);
  for (int i = 0; i < m; ++i) {
    RSFG_CHECK(rows[i] >= 0 && rows[i] < nrows, RSFG_EINVAL, "rsfg_entries: row out of range");
    // rows past the shear rows address the normal (In) rows, patch by patch
    rr[i] = rows[i] < nshear ? shear_row(h, rows[i]) : (rows[i] - nshear) * h->G + h->ncomp;
  }
  for (int j = 0; j < k; ++j) {
    RSFG_CHECK(cols[j] >= 0 && cols[j] < h->cols, RSFG_EINVAL, "rsfg_entries: col out of range");
    cc[j] = cols[j];
  }
  RSFG_HIP(hipSetDevice(h->device));
  const size_t ib = ((sizeof(int) * (m + k) + 7) / 8) * 8;
  int rc = ensure_tmp(h, ib + sizeof(double) * (size_t)m * k);
  if (rc) return rc;
  int* d_r = reinterpret_cast<int*>(h->d_tmp);
  int* d_c = d_r + m;
  double* d_o = reinterpret_cast<double*>(reinterpret_cast<char*>(h->d_tmp) + ib);
  RSFG_HIP(hipMemcpyAsync(d_r, rr.data(), sizeof(int) * m, hipMemcpyHostToDevice, h->stream));
  RSFG_HIP(hipMemcpyAsync(d_c, cc.data(), sizeof(int) * k, hipMemcpyHostToDevice, h->stream));
  const long tot = (long)m * k;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream,
                     h->K64, (long)h->ld, m, d_r, k, d_c, d_o);
  RSFG_HIP(hipGetLastError());
  RSFG_HIP(hipMemcpyAsync(out, d_o, sizeof(double) * tot, hipMemcpyDeviceToHost, h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  return RSFG_OK;
}

int rsfg_set_triangle_tolerance(rsfg_handle* h, double vertical_tol) {
  RSFG_CHECK(h, RSFG_EINVAL, "rsfg_set_triangle_tolerance: null handle");
  RSFG_CHECK(vertical_tol < 0.1, RSFG_EINVAL,
             "rsfg_set_triangle_tolerance: tolerance must be below 0.1 rad");
  h->tri_vtol = vertical_tol > 0.0 ? vertical_tol : 5.0e-15;
  return RSFG_OK;
}

int rsfg_traction_pairs(rsfg_handle* h, long np, const int* recv, const int* src, int full,
                        double* out) {
  RSFG_CHECK(h && np >= 0 && recv && src && out, RSFG_EINVAL, "rsfg_traction_pairs: bad args");
  if (np == 0) return RSFG_OK;
  for (long p = 0; p < np; ++p)
    RSFG_CHECK(recv[p] >= 0 && recv[p] < h->n && src[p] >= 0 && src[p] < h->n, RSFG_EINVAL,
               "rsfg_traction_pairs: index out of range");
  RSFG_HIP(hipSetDevice(h->device));
  const size_t ib = ((sizeof(int) * 2 * np + 7) / 8) * 8;
  int rc = ensure_tmp(h, ib + sizeof(double) * 6 * np);
  if (rc) return rc;
  int* d_r = reinterpret_cast<int*>(h->d_tmp);
  int* d_s = d_r + np;
  double* d_o = reinterpret_cast<double*>(reinterpret_cast<char*>(h->d_tmp) + ib);
  RSFG_HIP(hipMemcpyAsync(d_r, recv, sizeof(int) * np, hipMemcpyHostToDevice, h->stream));
  RSFG_HIP(hipMemcpyAsync(d_s, src, sizeof(int) * np, hipMemcpyHostToDevice, h->stream));
  rc = launch_pairs(h, np, d_r, d_s, full, d_o);
  if (rc) return rc;
  RSFG_HIP(hipMemcpyAsync(out, d_o, sizeof(double) * 6 * np, hipMemcpyDeviceToHost, h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  return RSFG_OK;
}

int rsfg_matvec_dev(rsfg_handle* h, int prec, const void* x_dev, void* y_dev) {
  RSFG_CHECK(h && x_dev && y_dev, RSFG_EINVAL, "rsfg_matvec_dev: null argument");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_matvec_dev: precision not assembled");
  RSFG_HIP(hipSetDevice(h->device));
  // input staging needs the padded buffer: use the scratch
  const size_t need = h->vin_elems ? h->vin_elems : vin_elems_for(h->cols);
  const size_t ybytes = sizeof(double) * h->rows_all;
  int rc = ensure_tmp(h, need * sizeof(double) + ybytes);
  if (rc) return rc;
  void* xin = h->d_tmp;
  double* yall = reinterpret_cast<double*>(reinterpret_cast<char*>(h->d_tmp) + need * sizeof(double));
  const size_t saved = h->vin_elems;
  h->vin_elems = need;
  rc = stage_input(h, prec, static_cast<const double*>(x_dev), xin);
  h->vin_elems = saved;
  if (rc) return rc;
  rc = matvec_any(h, prec, xin, yall);
  if (rc) return rc;
  if (h->G == h->ncomp) {
    RSFG_HIP(hipMemcpyAsync(y_dev, yall, ybytes, hipMemcpyDeviceToDevice, h->stream));
  } else {
    const int nr = h->n * h->ncomp;  // drop the normal rows
    hipLaunchKernelGGL(k_pick_shear, dim3((nr + 255) / 256), dim3(256), 0, h->stream, yall, h->n,
                       h->ncomp, h->G, static_cast<double*>(y_dev));
    RSFG_HIP(hipGetLastError());
  }
  return RSFG_OK;
}

int rsfg_matvec_dev_noalloc(rsfg_handle* h, int prec, const void* x_dev, void* y_dev) {
  RSFG_CHECK(h && x_dev && y_dev, RSFG_EINVAL, "rsfg_matvec_dev_noalloc: null argument");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_matvec_dev_noalloc: precision not assembled");
  const size_t need = vin_elems_for(h->cols);
  const size_t yall = (h->G == h->ncomp) ? 0 : (size_t)h->rows_all;
  RSFG_CHECK(need + yall <= kScratchDoubles, RSFG_EINVAL,
             "rsfg_matvec_dev_noalloc: operator too large for the static staging buffer ("
             "padded columns " + std::to_string(need) + " + rows " + std::to_string(yall) +
                 " > " + std::to_string(kScratchDoubles) + " doubles); use rsfg_matvec_dev");
  RSFG_HIP(hipSetDevice(h->device));
  double* sp = nullptr;
  int rc = scratch_ptr(h->device, &sp);
  if (rc) return rc;
  // the staging kernel writes every padded input element the product reads,
  // so nothing of an earlier call's contents is used
  const size_t saved = h->vin_elems;
  h->vin_elems = need;
  rc = stage_input(h, prec, static_cast<const double*>(x_dev), sp);
  h->vin_elems = saved;
  if (rc) return rc;
  if (yall == 0) return matvec_any(h, prec, sp, static_cast<double*>(y_dev));
  double* ya = sp + need;
  rc = matvec_any(h, prec, sp, ya);
  if (rc) return rc;
  const int nr = h->n * h->ncomp;  // drop the normal rows
  hipLaunchKernelGGL(k_pick_shear, dim3((nr + 255) / 256), dim3(256), 0, h->stream, ya, h->n,
                     h->ncomp, h->G, static_cast<double*>(y_dev));
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

int rsfg_matvec(rsfg_handle* h, int prec, const double* x, double* y) {
  RSFG_CHECK(h && x && y, RSFG_EINVAL, "rsfg_matvec: null argument");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_matvec: precision not assembled");
  RSFG_HIP(hipSetDevice(h->device));
  double *dx = nullptr, *dy = nullptr;
  const size_t nr = (size_t)h->ncomp * h->n;
  RSFG_HIP(hipMalloc(&dx, sizeof(double) * h->cols));
  RSFG_HIP(hipMalloc(&dy, sizeof(double) * nr));
  int rc = RSFG_OK;
  hipError_t e = hipMemcpyAsync(dx, x, sizeof(double) * h->cols, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) rc = hip_fail(e, "matvec upload", __FILE__, __LINE__);
  if (!rc) rc = rsfg_matvec_dev(h, prec, dx, dy);
  if (!rc) {
    e = hipMemcpyAsync(y, dy, sizeof(double) * nr, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = hip_fail(e, "matvec download", __FILE__, __LINE__);
  }
  hipFree(dx);
  hipFree(dy);
  return rc;
}

int rsfg_matvec_t(rsfg_handle* h, int prec, const double* x, double* y) {
  RSFG_CHECK(h && x && y, RSFG_EINVAL, "rsfg_matvec_t: null argument");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_matvec_t: precision not assembled");
  RSFG_CHECK(h->G == h->ncomp, RSFG_EINVAL, "rsfg_matvec_t: operator holds normal rows");
  RSFG_CHECK(prec != RSFG_PREC_BF16NF, RSFG_EINVAL,
             "rsfg_matvec_t: not available for the near-field bf16 format");
  RSFG_HIP(hipSetDevice(h->device));
  const int rows = h->rows_all, cols = h->cols;
  const int nsplit = std::max(1, std::min(64, rows / 64));
  double *dx = nullptr, *dy = nullptr, *dp = nullptr;
  RSFG_HIP(hipMalloc(&dx, sizeof(double) * rows));
  RSFG_HIP(hipMalloc(&dy, sizeof(double) * cols));
  RSFG_HIP(hipMalloc(&dp, sizeof(double) * cols * (size_t)nsplit));
  int rc = RSFG_OK;
  hipError_t e = hipMemcpyAsync(dx, x, sizeof(double) * rows, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) rc = hip_fail(e, "matvec_t upload", __FILE__, __LINE__);
  if (!rc) {
    switch (prec) {
      case RSFG_PREC_FP64: rc = matvec_t_launch<double>(h, h->K64, dx, dy, dp, nsplit); break;
      case RSFG_PREC_FP32: rc = matvec_t_launch<float>(h, h->K32, dx, dy, dp, nsplit); break;
      case RSFG_PREC_BF16X2: rc = matvec_t_launch<bf16x2_t>(h, h->Kbx2, dx, dy, dp, nsplit); break;
      default: rc = matvec_t_launch<bf16_t>(h, h->Kb16, dx, dy, dp, nsplit); break;
    }
  }
  if (!rc) {
    e = hipMemcpyAsync(y, dy, sizeof(double) * cols, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = hip_fail(e, "matvec_t download", __FILE__, __LINE__);
  }
  hipFree(dx);
  hipFree(dy);
  hipFree(dp);
  return rc;
}

int rsfg_matvec_bench(rsfg_handle* h, int prec, int nrep, double* ms) {
  RSFG_CHECK(h && nrep > 0 && ms, RSFG_EINVAL, "rsfg_matvec_bench: bad args");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_matvec_bench: precision not assembled");
  RSFG_HIP(hipSetDevice(h->device));
  const size_t need = vin_elems_for(h->cols);
  int rc = ensure_tmp(h, need * sizeof(double) + sizeof(double) * h->rows_all);
  if (rc) return rc;
  void* xin = h->d_tmp;
  double* y = reinterpret_cast<double*>(reinterpret_cast<char*>(h->d_tmp) + need * sizeof(double));
  RSFG_HIP(hipMemsetAsync(xin, 0, need * sizeof(double), h->stream));
  rc = matvec_any(h, prec, xin, y);  // warm
  if (rc) return rc;
  RSFG_HIP(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < nrep; ++i) {
    rc = matvec_any(h, prec, xin, y);
    if (rc) return rc;
  }
  RSFG_HIP(hipEventRecord(h->ev1, h->stream));
  RSFG_HIP(hipEventSynchronize(h->ev1));
  float t = 0.f;
  RSFG_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t / nrep;
  return RSFG_OK;
}

// Plain read stream: each wave reads a contiguous span with 16-byte
// nontemporal loads, eight in flight, and sums them (stored only if the sum
// hits an impossible value, so the loads stay).
__global__ __launch_bounds__(512) void k_read_stream(const v4f* __restrict__ a, size_t n4,
                                                    size_t per_wave, float* out) {
  constexpr int U = 8;
  const size_t wave = (size_t)blockIdx.x * 8 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const size_t b = min(wave * per_wave, n4), e = min(b + per_wave, n4);
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  size_t i = b + lane;
  for (; i + 64 * (U - 1) < e; i += 64 * U) {
    v4f t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = __builtin_nontemporal_load(a + i + 64 * u);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += t[u];
  }
  for (; i < e; i += 64) acc += __builtin_nontemporal_load(a + i);
  if (acc.x + acc.y + acc.z + acc.w == -1.0f) out[lane] = acc.x;
}

int rsfg_read_bench(rsfg_handle* h, double bytes, int nrep, double* ms) {
  RSFG_CHECK(h && nrep > 0 && ms && bytes >= 1 << 20 && bytes <= 64.0 * (1ull << 30),
             RSFG_EINVAL, "rsfg_read_bench: bad args (1 MiB <= bytes <= 64 GiB)");
  RSFG_HIP(hipSetDevice(h->device));
  const size_t n4 = (size_t)(bytes / 16.0);
  void* buf = nullptr;
  RSFG_HIP(hipMalloc(&buf, n4 * 16 + 256));
  float* out = reinterpret_cast<float*>(static_cast<char*>(buf) + n4 * 16);
  const int ncu = h->ncu > 0 ? h->ncu : 256;
  const int nwg = 2 * ncu;
  const size_t waves = (size_t)nwg * 8, per = (n4 + waves - 1) / waves;
  hipError_t e = hipMemsetAsync(buf, 0, n4 * 16 + 256, h->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_read_stream, dim3(nwg), dim3(512), 0, h->stream,
                       static_cast<const v4f*>(buf), n4, per, out);   // warm
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
  for (int k = 0; k < nrep && e == hipSuccess; ++k) {
    hipLaunchKernelGGL(k_read_stream, dim3(nwg), dim3(512), 0, h->stream,
                       static_cast<const v4f*>(buf), n4, per, out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, h->ev0, h->ev1);
  const hipError_t ef = hipFree(buf);
  if (e != hipSuccess) return hip_fail(e, "rsfg_read_bench", __FILE__, __LINE__);
  if (ef != hipSuccess) return hip_fail(ef, "rsfg_read_bench: free", __FILE__, __LINE__);
  *ms = t / nrep;
  return RSFG_OK;
}

int rsfg_set_friction(rsfg_handle* h, const rsfg_friction* f, const double* a, const double* b,
                      const double* dc) {
  RSFG_CHECK(h && f && a && b && dc, RSFG_EINVAL, "rsfg_set_friction: null argument");
  RSFG_CHECK(h->assembled, RSFG_ESTATE, "rsfg_set_friction: assemble first");
  RSFG_CHECK(f->law >= 1 && f->law <= 5, RSFG_EINVAL, "rsfg_set_friction: law must be 1..5");
  RSFG_CHECK(!f->calc_sigma_dot || h->with_normal, RSFG_EINVAL,
             "rsfg_set_friction: calc_sigma_dot needs the normal rows (with_normal)");
  for (int i = 0; i < h->n; ++i)
    RSFG_CHECK(a[i] > 0 && dc[i] > 0, RSFG_EINVAL, "rsfg_set_friction: a and D_c must be > 0");
  RSFG_HIP(hipSetDevice(h->device));
  FricParams& p = h->fp;
  p.v0 = f->v0;
  p.f0 = f->f0;
  p.eta = f->eta;
  p.vmin_state = f->vmin_state > 0 ? f->vmin_state : 1e-16;
  p.sato_beta = f->sato_beta > 0 ? f->sato_beta : 1e-2;
  p.kt_vc = f->kt_vc > 0 ? f->kt_vc : 1e-2 * f->v0;
  p.min_sigma = f->min_sigma;
  p.max_sigma = f->max_sigma;
  p.law = f->law;
  p.calc_sigma_dot = f->calc_sigma_dot;
  p.limit_sigma = f->limit_sigma;
  p.ncomp = h->ncomp;
  const int n = h->n;
  std::vector<double> c(3 * (size_t)n);
  memcpy(c.data(), a, sizeof(double) * n);
  memcpy(c.data() + n, b, sizeof(double) * n);
  memcpy(c.data() + 2 * n, dc, sizeof(double) * n);
  if (!h->d_coef) RSFG_HIP(hipMalloc(&h->d_coef, sizeof(double) * 3 * n));
  RSFG_HIP(hipMemcpy(h->d_coef, c.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  // loading block: 2 (ncomp + 1) n rates + n plate rates, ncomp <= 2
  if (!h->d_sinc) RSFG_HIP(hipMalloc(&h->d_sinc, sizeof(double) * 7 * n));
  h->have_friction = true;
  return alloc_state(h);
}

int rsfg_set_loading(rsfg_handle* h, double vpl, const double* vpl_patch, const double* taudot_add,
                     const double* sigdot_add) {
  RSFG_CHECK(h && h->have_friction, RSFG_ESTATE, "rsfg_set_loading: set friction first");
  RSFG_HIP(hipSetDevice(h->device));
  h->cont_ok = false;
  const int n = h->n, nc = h->ncomp;
  // backslip: slip rate -vpl along the slip direction (strike component in
  // the two-component mode), src/rsf_solve.c:268-331
  std::vector<double> x((size_t)h->cols, 0.0), y((size_t)h->rows_all, 0.0);
  for (int p = 0; p < n; ++p) x[(size_t)nc * p] = -(vpl_patch ? vpl_patch[p] : vpl);
  double *dx = nullptr, *dy = nullptr;
  RSFG_HIP(hipMalloc(&dx, sizeof(double) * h->cols));
  RSFG_HIP(hipMalloc(&dy, sizeof(double) * h->rows_all));
  RSFG_HIP(hipMemcpy(dx, x.data(), sizeof(double) * h->cols, hipMemcpyHostToDevice));
  const size_t need = vin_elems_for(h->cols);
  int rc = ensure_tmp(h, need * sizeof(double));
  if (!rc) rc = stage_input(h, RSFG_PREC_FP64, dx, h->d_tmp);
  if (!rc) rc = matvec_any(h, RSFG_PREC_FP64, h->d_tmp, dy);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(y.data(), dy, sizeof(double) * h->rows_all, hipMemcpyDeviceToHost,
                              h->stream);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(dx);
  hipFree(dy);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "rsfg_set_loading", __FILE__, __LINE__);
  // [full shear | full normal] = backslip products + additive rates;
  // [additive shear | additive normal] and the plate rate per patch serve
  // the reduced-precision paths, which fold the backslip into K (v - vpl)
  const size_t nl = (size_t)(nc + 1) * n;
  std::vector<double> s(2 * nl + n, 0.0);
  for (int p = 0; p < n; ++p) {
    const double tadd = taudot_add ? taudot_add[p] : 0.0;
    const double sadd = (sigdot_add && h->fp.calc_sigma_dot) ? sigdot_add[p] : 0.0;
    for (int k = 0; k < nc; ++k) s[(size_t)nc * p + k] = y[(size_t)p * h->G + k];
    s[(size_t)nc * p] += tadd;
    s[nl + (size_t)nc * p] = tadd;
    const double sn = (h->G > nc) ? y[(size_t)p * h->G + nc] : 0.0;
    s[(size_t)nc * n + p] = sn + sadd;
    s[nl + (size_t)nc * n + p] = sadd;
    s[2 * nl + p] = vpl_patch ? vpl_patch[p] : vpl;
  }
  RSFG_HIP(hipMemcpy(h->d_sinc, s.data(), sizeof(double) * s.size(), hipMemcpyHostToDevice));
  h->have_loading = true;
  return RSFG_OK;
}

int rsfg_set_state(rsfg_handle* h, double t, const double* X) {
  RSFG_CHECK(h && X && h->have_friction, RSFG_ESTATE, "rsfg_set_state: set friction first");
  RSFG_HIP(hipSetDevice(h->device));
  const int n = h->n, D = h->dim;
  std::vector<double> soa((size_t)D * n);
  for (int p = 0; p < n; ++p)
    for (int f = 0; f < D; ++f) soa[(size_t)f * n + p] = X[(size_t)p * D + f];
  h->cont_ok = false;
  h->xcur = 0;
  RSFG_HIP(hipMemcpy(h->Xb[0], soa.data(), sizeof(double) * D * n, hipMemcpyHostToDevice));
  h->t = t;
  h->have_state = true;
  return RSFG_OK;
}

int rsfg_get_state(rsfg_handle* h, double* t, double* X) {
  RSFG_CHECK(h && X && h->have_state, RSFG_ESTATE, "rsfg_get_state: no state");
  RSFG_HIP(hipSetDevice(h->device));
  const int n = h->n, D = h->dim;
  std::vector<double> soa((size_t)D * n);
  RSFG_HIP(hipMemcpy(soa.data(), h->Xb[h->xcur], sizeof(double) * D * n, hipMemcpyDeviceToHost));
  for (int p = 0; p < n; ++p)
    for (int f = 0; f < D; ++f) X[(size_t)p * D + f] = soa[(size_t)f * n + p];
  if (t) *t = h->t;
  return RSFG_OK;
}

int rsfg_set_step_log(rsfg_handle* h, long long capacity) {
  RSFG_CHECK(h && capacity >= 0, RSFG_EINVAL, "rsfg_set_step_log: bad arguments");
  RSFG_HIP(hipSetDevice(h->device));
  if (h->stream) RSFG_HIP(hipStreamSynchronize(h->stream));
  hipFree(h->d_hlog);
  h->d_hlog = nullptr;
  h->hlog_cap = 0;
  if (capacity > 0) {
    RSFG_HIP(hipMalloc(&h->d_hlog, sizeof(double) * capacity));
    h->hlog_cap = capacity;
  }
  return RSFG_OK;
}

int rsfg_get_step_log(rsfg_handle* h, double* steps, long long capacity, long long* nsteps) {
  RSFG_CHECK(h && nsteps && capacity >= 0, RSFG_EINVAL, "rsfg_get_step_log: bad arguments");
  RSFG_CHECK(h->d_hlog != nullptr, RSFG_ESTATE, "rsfg_get_step_log: no step log (rsfg_set_step_log)");
  RSFG_HIP(hipSetDevice(h->device));
  const long long done = h->h_ctl ? h->h_ctl->steps : 0;
  RSFG_CHECK(done <= h->hlog_cap, RSFG_ESTATE,
             "rsfg_get_step_log: the last solve took more steps than the log holds");
  *nsteps = done;
  const long long k = std::min(done, capacity);
  if (steps && k > 0)
    RSFG_HIP(hipMemcpy(steps, h->d_hlog, sizeof(double) * k, hipMemcpyDeviceToHost));
  return RSFG_OK;
}

int rsfg_read_table(const char* fname, int n, int ncols, double* out) {
  RSFG_CHECK(fname && out && n > 0 && ncols > 0, RSFG_EINVAL, "rsfg_read_table: bad arguments");
  FILE* in = fopen(fname, "r");
  RSFG_CHECK(in != nullptr, RSFG_EINVAL, std::string("rsfg_read_table: cannot open ") + fname);
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < ncols; ++c) {
      if (fscanf(in, "%lf", out + (size_t)i * ncols + c) != 1) {
        fclose(in);
        set_error("rsfg_read_table: error reading row " + std::to_string(i) + " column " +
                  std::to_string(c) + " of " + fname);
        return RSFG_EINVAL;
      }
    }
  }
  fclose(in);
  return RSFG_OK;
}

// PETSc binary vectors: int32 class id 1211214, int32 length, then the
// values as IEEE doubles, all big endian
constexpr int32_t kVecClassId = 1211214;
constexpr double kCkptMagic = 20260726.0;
constexpr int kCkptMeta = 16;

static void put_be32(std::vector<unsigned char>& b, int32_t v) {
  const uint32_t u = (uint32_t)v;
  for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(u >> (8 * k)));
}
static void put_be_double(std::vector<unsigned char>& b, double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  for (int k = 7; k >= 0; --k) b.push_back((unsigned char)(u >> (8 * k)));
}
static bool get_be32(FILE* f, int32_t* v) {
  unsigned char c[4];
  if (fread(c, 1, 4, f) != 4) return false;
  *v = (int32_t)(((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | c[3]);
  return true;
}
static bool get_be_vec(FILE* f, std::vector<double>& v) {
  int32_t id, len;
  if (!get_be32(f, &id) || !get_be32(f, &len) || id != kVecClassId || len < 0) return false;
  std::vector<unsigned char> c((size_t)len * 8);
  if (fread(c.data(), 1, c.size(), f) != c.size()) return false;
  v.resize(len);
  for (int32_t i = 0; i < len; ++i) {
    uint64_t u = 0;
    for (int k = 0; k < 8; ++k) u = (u << 8) | c[(size_t)i * 8 + k];
    memcpy(&v[i], &u, 8);
  }
  return true;
}

int rsfg_write_checkpoint(rsfg_handle* h, const char* fname, int slip_mode, int law, long step,
                          double dt) {
  RSFG_CHECK(h && fname, RSFG_EINVAL, "rsfg_write_checkpoint: null argument");
  const int n = h->n, D = h->dim;
  std::vector<double> X((size_t)D * n);
  double t = 0.0;
  int rc = rsfg_get_state(h, &t, X.data());
  if (rc) return rc;
  RSFG_CHECK((size_t)D * n <= 0x7fffffff, RSFG_EINVAL, "rsfg_write_checkpoint: state too large");
  double meta[kCkptMeta] = {0.0};
  meta[0] = kCkptMagic;
  meta[1] = 1.0;
  meta[2] = n;
  meta[3] = D;
  meta[4] = slip_mode;
  meta[5] = law;
  meta[6] = (double)step;
  meta[7] = t;
  meta[8] = dt;
  std::vector<unsigned char> b;
  b.reserve(16 + 8 * (kCkptMeta + X.size()));
  put_be32(b, kVecClassId);
  put_be32(b, kCkptMeta);
  for (double m : meta) put_be_double(b, m);
  put_be32(b, kVecClassId);
  put_be32(b, (int32_t)X.size());
  for (double x : X) put_be_double(b, x);
  const std::string tmp = std::string(fname) + ".tmp", prev = std::string(fname) + ".prev";
  FILE* f = fopen(tmp.c_str(), "wb");
  RSFG_CHECK(f != nullptr, RSFG_EINVAL, "rsfg_write_checkpoint: cannot create " + tmp);
  const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
  const bool closed = fclose(f) == 0;
  RSFG_CHECK(ok && closed, RSFG_EINVAL, "rsfg_write_checkpoint: write to " + tmp + " failed");
  rename(fname, prev.c_str());  // fails harmlessly on the first write
  RSFG_CHECK(rename(tmp.c_str(), fname) == 0, RSFG_EINVAL,
             std::string("rsfg_write_checkpoint: rename to ") + fname + " failed");
  return RSFG_OK;
}

int rsfg_read_checkpoint(rsfg_handle* h, const char* fname, int* slip_mode, int* law,
                         long* step, double* t, double* dt) {
  RSFG_CHECK(h && fname, RSFG_EINVAL, "rsfg_read_checkpoint: null argument");
  FILE* f = fopen(fname, "rb");
  RSFG_CHECK(f != nullptr, RSFG_EINVAL, std::string("rsfg_read_checkpoint: cannot open ") + fname);
  std::vector<double> meta, X;
  const bool ok = get_be_vec(f, meta) && get_be_vec(f, X);
  fclose(f);
  RSFG_CHECK(ok && meta.size() == (size_t)kCkptMeta && meta[0] == kCkptMagic, RSFG_EINVAL,
             std::string("rsfg_read_checkpoint: ") + fname + " is not an rsf_solve checkpoint");
  RSFG_CHECK((int)meta[2] == h->n, RSFG_EINVAL,
             "rsfg_read_checkpoint: checkpoint has " + std::to_string((long)meta[2]) +
                 " patches, geometry has " + std::to_string(h->n));
  RSFG_CHECK((int)meta[3] == h->dim && X.size() == (size_t)h->dim * h->n, RSFG_EINVAL,
             "rsfg_read_checkpoint: checkpoint dim " + std::to_string((long)meta[3]) +
                 ", run dim " + std::to_string(h->dim));
  const int rc = rsfg_set_state(h, meta[7], X.data());
  if (rc) return rc;
  if (slip_mode) *slip_mode = (int)meta[4];
  if (law) *law = (int)meta[5];
  if (step) *step = (long)meta[6];
  if (t) *t = meta[7];
  if (dt) *dt = meta[8];
  return RSFG_OK;
}

int rsfg_rhs(rsfg_handle* h, int prec, const double* X, double* F, int* domain_ok) {
  RSFG_CHECK(h && X && F, RSFG_EINVAL, "rsfg_rhs: null argument");
  RSFG_CHECK(h->have_friction && h->have_loading, RSFG_ESTATE,
             "rsfg_rhs: set friction and loading first");
  RSFG_CHECK(prec_ok(h, prec), RSFG_ESTATE, "rsfg_rhs: precision not assembled");
  RSFG_HIP(hipSetDevice(h->device));
  h->cont_ok = false;
  const int n = h->n, D = h->dim, nc = h->ncomp;
  std::vector<double> soa((size_t)D * n);
  for (int p = 0; p < n; ++p)
    for (int f = 0; f < D; ++f) soa[(size_t)f * n + p] = X[(size_t)p * D + f];
  double* Y = h->Yb[0];
  double* Fd = h->Yb[1];
  int rc = ensure_tmp(h, sizeof(int) * n);
  if (rc) return rc;
  int* okd = reinterpret_cast<int*>(h->d_tmp);
  RSFG_HIP(hipMemcpyAsync(Y, soa.data(), sizeof(double) * D * n, hipMemcpyHostToDevice, h->stream));
  RSFG_HIP(hipMemsetAsync(h->vin[0], 0, sizeof(double) * h->vin_elems, h->stream));
  const dim3 gv((n + 255) / 256);
  const bool f64 = (prec == RSFG_PREC_FP64);
#define RSFG_VEL(VT)                                                                          \
  if (nc == 1)                                                                                \
    hipLaunchKernelGGL((k_vel<VT, 1>), gv, dim3(256), 0, h->stream, Y, h->d_coef, h->fp, n,  \
                       reinterpret_cast<VT*>(h->vin[0]), okd);                                \
  else                                                                                        \
    hipLaunchKernelGGL((k_vel<VT, 2>), gv, dim3(256), 0, h->stream, Y, h->d_coef, h->fp, n,  \
                       reinterpret_cast<VT*>(h->vin[0]), okd);
  if (f64) {
    RSFG_VEL(double)
  } else {
    RSFG_VEL(float)
  }
#undef RSFG_VEL
  const int G = h->G;
  const long ld = h->ld;
  const int rows = h->rows_all, cols = h->cols;
#define RSFG_RHS(TK, NC, GG)                                                                  \
  {                                                                                           \
    hipLaunchKernelGGL((k_rhs<TK, NC, GG>), dim3((n + 3) / 4), dim3(256), 0,                  \
                       h->stream, static_cast<const TK*>(operator_ptr(h, prec)), ld, rows,    \
                       cols, reinterpret_cast<const typename KTraits<TK>::VT*>(h->vin[0]), Y, \
                       h->d_coef, h->d_sinc, h->fp, n, Fd);                                   \
  }
#define RSFG_RHS_T(TK)                              \
  if (nc == 1 && G == 1) RSFG_RHS(TK, 1, 1)         \
  else if (nc == 1) RSFG_RHS(TK, 1, 2)              \
  else if (G == 2) RSFG_RHS(TK, 2, 2)               \
  else RSFG_RHS(TK, 2, 3)
  switch (prec) {
    case RSFG_PREC_FP64: RSFG_RHS_T(double) break;
    case RSFG_PREC_FP32: RSFG_RHS_T(float) break;
    case RSFG_PREC_BF16X2: RSFG_RHS_T(bf16x2_t) break;
    default: RSFG_RHS_T(bf16_t) break;
  }
#undef RSFG_RHS_T
#undef RSFG_RHS
  RSFG_HIP(hipGetLastError());
  std::vector<int> okh(n);
  RSFG_HIP(hipMemcpyAsync(soa.data(), Fd, sizeof(double) * D * n, hipMemcpyDeviceToHost, h->stream));
  RSFG_HIP(hipMemcpyAsync(okh.data(), okd, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  for (int p = 0; p < n; ++p)
    for (int f = 0; f < D; ++f) F[(size_t)p * D + f] = soa[(size_t)f * n + p];
  if (domain_ok) {
    int ok = 1;
    for (int p = 0; p < n; ++p) ok &= okh[p];
    *domain_ok = ok;
  }
  return RSFG_OK;
}

int rsfg_solver_defaults(rsfg_solver_opts* o) {
  RSFG_CHECK(o, RSFG_EINVAL, "rsfg_solver_defaults: null");
  memset(o, 0, sizeof(*o));
  o->method = RSFG_RK5DP;
  o->prec = RSFG_PREC_FP64;
  o->rtol = 1e-4;
  o->atol[0] = 1e-12;  // psi (RSF_AERROR_PSI)
  for (int f = 1; f < 6; ++f) o->atol[f] = 1e-3;  // tau, sigma [Pa]; slip [m]
  o->dt_init = 1.0;
  o->dt_max = 1e10;
  o->dt_monitor = 5.0 * 31557600.0;
  o->rdx_monitor = 1e-4;
  o->adx_monitor = 0.0;
  o->vel_event = 1e-3;
  o->vel_event_hyst = 0.5;
  o->rupture_vth = 1e-3;
  o->max_steps = 1000000000L;
  o->max_domain_reject = 5000;
  o->steps_per_launch = 32;
  return RSFG_OK;
}

int rsfg_solve(rsfg_handle* h, const rsfg_solver_opts* o, double t_end, rsfg_stats* st) {
  RSFG_CHECK(h && o, RSFG_EINVAL, "rsfg_solve: null argument");
  RSFG_CHECK(h->have_state && h->have_loading, RSFG_ESTATE,
             "rsfg_solve: set friction, loading and state first");
  RSFG_HIP(hipSetDevice(h->device));
  if (h->xcur != 0) {  // the solver starts from buffer 0
    RSFG_HIP(hipMemcpyAsync(h->Xb[0], h->Xb[h->xcur], sizeof(double) * h->dim * h->n,
                            hipMemcpyDeviceToDevice, h->stream));
    h->xcur = 0;
  }
  return solve_device(h, o, t_end, st, false);
}

int rsfg_solve_continue(rsfg_handle* h, const rsfg_solver_opts* o, double t_end,
                        rsfg_stats* st) {
  RSFG_CHECK(h && o, RSFG_EINVAL, "rsfg_solve_continue: null argument");
  RSFG_HIP(hipSetDevice(h->device));
  return solve_device(h, o, t_end, st, true);
}

int rsfg_stage_bench(rsfg_handle* h, int prec, int nrep, double* ms) {
  RSFG_CHECK(h && nrep > 0 && ms, RSFG_EINVAL, "rsfg_stage_bench: bad args");
  RSFG_CHECK(h->have_state && h->have_loading, RSFG_ESTATE,
             "rsfg_stage_bench: set friction, loading and state first");
  RSFG_HIP(hipSetDevice(h->device));
  if (h->xcur != 0) {
    RSFG_HIP(hipMemcpyAsync(h->Xb[0], h->Xb[h->xcur], sizeof(double) * h->dim * h->n,
                            hipMemcpyDeviceToDevice, h->stream));
    h->xcur = 0;
  }
  return stage_bench(h, prec, nrep, ms);
}

int rsfg_events(rsfg_handle* h, double* rows, int* n) {
  RSFG_CHECK(h && rows && n, RSFG_EINVAL, "rsfg_events: null argument");
  const int k = std::min(*n, h->nevent);
  if (k > 0)
    RSFG_HIP(hipMemcpy(rows, h->d_ev, sizeof(double) * 8 * k, hipMemcpyDeviceToHost));
  *n = k;
  return RSFG_OK;
}

int rsfg_catalog(rsfg_handle* h, double* rows, int* n) {
  RSFG_CHECK(h && rows && n, RSFG_EINVAL, "rsfg_catalog: null argument");
  const int k = std::min(*n, h->ncat);
  if (k > 0)
    RSFG_HIP(hipMemcpy(rows, h->d_catr, sizeof(double) * kCatCols * k, hipMemcpyDeviceToHost));
  *n = k;
  return RSFG_OK;
}

int rsfg_monitor(rsfg_handle* h, double* rows, int* n) {
  RSFG_CHECK(h && rows && n, RSFG_EINVAL, "rsfg_monitor: null argument");
  const int k = std::min(*n, h->nmon);
  if (k > 0)
    RSFG_HIP(hipMemcpy(rows, h->d_mon, sizeof(double) * 10 * k, hipMemcpyDeviceToHost));
  *n = k;
  return RSFG_OK;
}

}  // extern "C"
