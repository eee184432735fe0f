// Device-resident adaptive Runge-Kutta cycle integration.
//
// One attempted step of an s-stage FSAL pair (rk5dp: s = 7, rk3bs: s = 4)
// is s + 1 launches:
//   k_prep            Y_1 = X + h a10 K_0, v(Y_1), stage domain check
//   k_stage(i), i=1..s-1
//                     y = K v(Y_i) streamed from HBM; epilogue per patch:
//                     K_i = F(Y_i, y) (rsf_ODE_RHSFunction), then either
//                     Y_{i+1} and v(Y_{i+1}) for the next launch, or (last
//                     stage) the weighted error and per-workgroup partial
//                     reductions of the new state
//   k_ctrl            one workgroup: grid reduction and the controller below.
// A separate control launch replaces a last-workgroup ticket on purpose: a
// device-scope fence in each of ~10^3 workgroups costs more than the whole
// stage's HBM stream on gfx950 (measured +50 us on a 40 us stage).
// The controller restates the PETSc behaviour the reference relies on
// (src/rsf_solve.c:440-546): TSAdapt basic (safety 0.9, x0.5 after a
// failed attempt, exponent -1/order, clip [0.5, 2], h <= min(dt_max,
// dt_monitor)), stage domain failure -> h * 0.25 (rsf_domain_check,
// src/rsf_engine.c:309-362), threshold events on log10 max|v| with
// hysteresis (src/rsf_solve.c:677-703, src/rsf_output.c:1000-1150)
// located on the step's dense output, and the change-triggered monitor
// (src/rsf_output.c:718-864).  The host only launches captured graphs of
// step attempts and polls the halted flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gemv_dev.h"
#include "rsf_dev.h"
#include "rsfg_common.h"
#include "stage_dev.h"

namespace rsfg {

// ------------------------------------------------------------ kernels
// prep: first stage of an attempt (also Xmon copy and rupture marking)
template <typename VT, int NC>
__global__ __launch_bounds__(256) void k_prep(StageArgs a) {
  prep_patch<VT, NC>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// initial velocity field (before K_0 = F(X_0))
template <typename VT, int NC>
__global__ __launch_bounds__(256) void k_init_vel(StageArgs a) {
  init_vel_patch<VT, NC>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// Wave w of workgroup b owns patch p = kStagePPB b + w: it streams the patch's G
// operator rows (ncomp shear rows, then the normal row when present) and
// leaves the row sums in LDS.  Wave 0 then runs the epilogue with one lane
// per patch of the workgroup: the epilogue is a long FP64 chain (exp, sinh,
// log), and running it on kStagePPB lanes of one wave instead of one lane in
// each of kStagePPB waves divides its share of the SIMDs' issue time by
// kStagePPB.  Stages that reduce (the initial evaluation and the final
// stage) combine the workgroup's patches with cross-lane shuffles and write
// one column of partials[kNRed][gridDim.x]; k_ctrl reduces them.
#ifndef RSFG_STAGE_PPB
#define RSFG_STAGE_PPB 8
#endif
constexpr int kStagePPB = RSFG_STAGE_PPB;  // patches (= waves) per workgroup
#ifndef RSFG_STAGE_MINW
#define RSFG_STAGE_MINW 4
#endif
static_assert(kStagePPB >= 1 && kStagePPB <= 16, "one wave per patch, at most 1024 threads");

// Epilogue operands of one patch whose addresses are known at launch: the
// state-only record of Y_s, the accepted state X, the tableau-weighted sum
// of the earlier stage derivatives and the stressing-rate loading; with
// kAll also the friction coefficients, the state itself (reducing stages)
// and the monitor snapshot (final stage).  The stage kernels request them
// before streaming their rows, so the epilogue starts on data already in
// registers.
template <int NC>
struct StageOperands {
  double P[PreDim<NC>::v], sn[NC], snn, x0[Dim<NC>::v], ws[Dim<NC>::v];
  PatchCoef cf;
  double y[Dim<NC>::v], xm[Dim<NC>::v];
};

template <int NC, bool kAll>
__device__ __forceinline__ void stage_operands(const StageArgs& a, const int* ks, size_t sst,
                                               const double* X, const double* Xn, int p,
                                               bool own, StageOperands<NC>& o) {
  constexpr int D = Dim<NC>::v;
  constexpr int PD = PreDim<NC>::v;
  const int n = a.n, S = a.tb.s, s = a.s;
  const bool mid = (s > 0) && (s < S - 1);
#pragma unroll
  for (int k = 0; k < PD; ++k) o.P[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NC; ++k) o.sn[k] = 0.0;
  o.snn = 0.0;
#pragma unroll
  for (int f = 0; f < D; ++f) o.x0[f] = o.ws[f] = o.y[f] = o.xm[f] = 0.0;
  o.cf.a = o.cf.b = o.cf.dc = 0.0;
  if (!own) return;
  if (s > 0) {
    const double* Ps = a.Pb[s & 1];
#pragma unroll
    for (int k = 0; k < PD; ++k) o.P[k] = Ps[k * n + p];
#pragma unroll
    for (int f = 0; f < D; ++f) o.x0[f] = X[f * n + p];
    // sum_j A[s+1][j] K_j over j < s (middle stages), or
    // sum_j bh[j] K_j over j < S - 1 (final stage)
    // (a fully unrolled loop: ks stays in registers, every index constant)
    const int nj = mid ? s : S - 1;
#pragma unroll
    for (int j = 0; j < kMaxStages; ++j) {
      const double w = (j < nj) ? (mid ? a.tb.A[s + 1][j] : a.tb.bh[j]) : 0.0;
      if (w != 0.0) {
        const double* Kj = a.Stg[0] + (size_t)ks[j] * sst;
#pragma unroll
        for (int f = 0; f < D; ++f) o.ws[f] += w * Kj[f * n + p];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) o.sn[k] = a.sinc[NC * p + k];
  o.snn = a.sinc[NC * n + p];
  if constexpr (kAll) {
    o.cf = load_coef(a.coef, n, p);
    if (!mid) {
      const double* Yc = (s == 0) ? X : Xn;
#pragma unroll
      for (int f = 0; f < D; ++f) o.y[f] = Yc[f * n + p];
    }
    if (s == S - 1) {
#pragma unroll
      for (int f = 0; f < D; ++f) o.xm[f] = a.Xmon[f * n + p];
    }
  }
}

__device__ __forceinline__ void r9_neutral(double* r9) {
  r9[RED_ERR] = 0.0; r9[RED_VMAX] = 0.0; r9[RED_VSUM] = 0.0; r9[RED_SLIP] = 0.0;
  r9[RED_MU] = 0.0; r9[RED_SMAX] = -1e30; r9[RED_SMIN] = 1e30; r9[RED_XN2] = 0.0;
  r9[RED_DX2] = 0.0;
}

// Stage epilogue of patch p from its G operator row sums rsum (shear rows,
// then the normal row when present): K_s = F(Y_s, rsum) into its stage
// slot, then Y_{s+1}, its state-only record and v(Y_{s+1}) (middle stages),
// or the patch's terms of the launch's reductions r9 (initial evaluation
// and final stage: error norm, slip-rate and stress extremes, ...).
template <typename VT, int NC, int G, bool kAll>
__device__ __forceinline__ void stage_epilogue(const StageArgs& a, DevCtl* c, int kss,
                                               size_t sst, double hstep, const double* X,
                                               double* Xn, int p, const double* rsum,
                                               StageOperands<NC>& o, double* r9) {
  constexpr int D = Dim<NC>::v;
  constexpr int PD = PreDim<NC>::v;
  const int n = a.n, S = a.tb.s, s = a.s;
  const bool reduce = (s == 0) || (s == S - 1);
  const bool mid = (s > 0) && (s < S - 1);
  const PatchCoef cf = kAll ? o.cf : load_coef(a.coef, n, p);
  double y[D], kf[D], td[NC];
  if (reduce) {
    const double* Yc = (s == 0) ? X : Xn;
#pragma unroll
    for (int f = 0; f < D; ++f) y[f] = kAll ? o.y[f] : Yc[f * n + p];
  }
  if (s == 0) rsf_patch_pre<NC>(a.fp, cf, y, o.P);
#pragma unroll
  for (int k = 0; k < NC; ++k) td[k] = rsum[k];
  const double sdn = (G > NC) ? rsum[NC] : 0.0;
  rsf_patch_finish<NC>(a.fp, o.P, td, sdn, o.sn, o.snn, kf);
  double* Ks = a.Stg[0] + (size_t)kss * sst;
#pragma unroll
  for (int f = 0; f < D; ++f) Ks[f * n + p] = kf[f];
  if (mid) {
    // next stage state Y_{s+1} = X + h (sum_{j<s} A[s+1][j] K_j + A[s+1][s] K_s)
    // and its state-only record
    // (only the last stage state is stored: it is the new solution; the
    // others live on in their state-only records)
    const double w = a.tb.A[s + 1][s];
    const bool last = (s + 1 == S - 1);
    double yn[D];
#pragma unroll
    for (int f = 0; f < D; ++f) {
      yn[f] = o.x0[f] + hstep * (o.ws[f] + w * kf[f]);
      if (last) Xn[f * n + p] = yn[f];
    }
    double Pn[PD];
    rsf_patch_pre<NC>(a.fp, cf, yn, Pn);
    store_pre<NC>(Pn, a.Pb[(s + 1) & 1], n, p);
    write_vel_pre<VT, NC>(Pn, static_cast<VT*>(a.vout), p, a.vpl);
    if (!rsf_patch_ok<NC>(a.fp, cf, yn)) atomicOr(&c->stage_fail, 1);
  } else {
    // reductions of the state y (initial state or the new solution)
    if (s == S - 1) {
      double en = 0.0;
#pragma unroll
      for (int f = 0; f < D; ++f) {
        const double e = o.ws[f] + a.tb.bh[S - 1] * kf[f];
        const double ye = o.x0[f] + hstep * e;
        const double tol = a.so.atol[f] + a.so.rtol * fmax(fabs(y[f]), fabs(ye));
        if (tol > 0.0) en = fmax(en, fabs(y[f] - ye) / tol);
        const double dm = (kAll ? o.xm[f] : a.Xmon[f * n + p]) - y[f];
        r9[RED_DX2] += dm * dm;
      }
      r9[RED_ERR] = en;
    }
    const double vm = pre_vmag<NC>(o.P);
    r9[RED_VMAX] = vm;
    r9[RED_VSUM] = vm;
    r9[RED_SLIP] = y[NC + 2];
    r9[RED_MU] = y[1] / y[NC + 1];
    r9[RED_SMAX] = y[NC + 1];
    r9[RED_SMIN] = y[NC + 1];
#pragma unroll
    for (int f = 0; f < D; ++f) r9[RED_XN2] += y[f] * y[f];
  }
}

// one wave's reductions into column col of partials[kNRed][ncol]; lanes
// without a patch hold the neutral values
__device__ __forceinline__ void write_partials(const StageArgs& a, const double* r9, int col,
                                               int ncol) {
#pragma unroll
  for (int k = 0; k < kNRed; ++k) {
    const double q = wave_reduce(r9[k], red_kind(k));
    if ((threadIdx.x & 63) == 0) a.partials[(size_t)k * ncol + col] = q;
  }
}

template <typename TK, int NC, int G>
__global__ __launch_bounds__(64 * kStagePPB, RSFG_STAGE_MINW) void k_stage(StageArgs a) {
  using VT = typename KTraits<TK>::VT;
  __shared__ double rs[kStagePPB][G];
  DevCtl* c = a.ctl;
  // The control fields the launch needs are read with the halted flag, in
  // one memory round trip (they share a few cache lines); the epilogue then
  // addresses its operands without a second dependent read.  Stage slots
  // are consecutive blocks of one allocation, so slot pointers are formed
  // by arithmetic instead of an indexed read of the argument block.
  const int halted = c->halted, xc = c->xcur;
  const double hstep = c->h;
  int ks[kMaxStages];
#pragma unroll
  for (int j = 0; j < kMaxStages; ++j) ks[j] = c->kslot[j];
  const int kss = c->kslot[a.s];   // this stage's derivative slot
  if (halted) return;
  const size_t sst = a.Stg[1] - a.Stg[0];
  const int n = a.n, S = a.tb.s, s = a.s;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = lane;
  const int p = blockIdx.x * kStagePPB + t;
  const double* X = xc ? a.Xb[1] : a.Xb[0];
  double* Xn = xc ? a.Xb[0] : a.Xb[1];
  const bool reduce = (s == 0) || (s == S - 1);
  // Wave 0 requests the epilogue operands before it streams its rows.
  // (Fetching the remaining operands as well was measured slower: the two
  // workgroups sharing a CU then stream in lockstep instead of one after
  // the other.)
  StageOperands<NC> op;
  stage_operands<NC, false>(a, ks, sst, X, Xn, p, wave == 0 && t < kStagePPB && p < n, op);
  {
    const int pw = min(blockIdx.x * kStagePPB + wave, n - 1);
    double rowsum[G];
    if constexpr (std::is_same<TK, bf16nf_t>::value) {
      const int o = a.nf_off[pw];
      gemv_nf_wave<G>(static_cast<const bf16_t*>(a.K), a.Knf_lo + (size_t)o * G * kNfChunk,
                      a.nf_list + o, a.nf_off[pw + 1] - o, a.ld, pw * G, a.rows_all, a.cols,
                      static_cast<const float*>(a.vin), rowsum);
    } else {
      gemv_wave<TK, G>(static_cast<const TK*>(a.K), a.ld, pw * G, a.rows_all, a.cols,
                       static_cast<const VT*>(a.vin), rowsum);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < G; ++k) rs[wave][k] = rowsum[k];
    }
  }
  __syncthreads();
  if (wave != 0) return;
  double r9[kNRed];
  r9_neutral(r9);
  if (t < kStagePPB && p < n) {
    double rsum[G];
#pragma unroll
    for (int k = 0; k < G; ++k) rsum[k] = rs[t][k];
    stage_epilogue<VT, NC, G, false>(a, c, kss, sst, hstep, X, Xn, p, rsum, op, r9);
  }
  if (reduce) write_partials(a, r9, blockIdx.x, gridDim.x);
}

// Persistent stage kernel of the FP32-input formats (FP32, bf16 hi+lo,
// bf16, bf16 near-field): one workgroup of kPW waves per CU, x copied into
// LDS once per launch.  Wave w of workgroup b owns the ppw consecutive
// patches starting at (kPW b + w) ppw (ppw <= kMaxPPW) and streams their G
// rows one patch after the other against the LDS copy of x.  The row sums
// of the workgroup's kPW ppw patches then meet in LDS and kEW waves, one
// per SIMD, run the epilogue with one lane per patch: the epilogue is a
// long FP64 chain at the end of the launch, when no stream is left to hide
// it, and one wave per SIMD halves its issue time against two.  The
// epilogue waves request their patches' operands before streaming.  Each
// epilogue wave writes its own column of the reductions.  Against the
// per-patch kernel this removes the x slices' L2 traffic and the
// per-workgroup launch, barrier and epilogue gaps; every sum is formed in
// the same order.
constexpr int kEW = 4;          // epilogue waves per workgroup (one per SIMD)
constexpr int kMaxPPW = 32;     // patches per wave: kPW kMaxPPW / kEW <= 64 lanes

template <typename TK, int NC, int G, bool CB>
__global__ __launch_bounds__(64 * kPW, 1) void k_stage_p(StageArgs a, int ppw) {
  using VT = typename KTraits<TK>::VT;
  using TH = typename std::conditional<std::is_same<TK, bf16nf_t>::value, bf16_t, TK>::type;
  static_assert(std::is_same<VT, float>::value, "FP32-input formats only");
  static_assert(kPW * kMaxPPW / kEW <= 64, "one lane per patch in the epilogue waves");
  __shared__ __attribute__((aligned(16))) float xs[kXL];
  __shared__ double red[kPW * kMaxPPW][G];
  DevCtl* c = a.ctl;
  const int halted = c->halted, xc = c->xcur;
  const double hstep = c->h;
  int ks[kMaxStages];
#pragma unroll
  for (int j = 0; j < kMaxStages; ++j) ks[j] = c->kslot[j];
  const int kss = c->kslot[a.s];   // this stage's derivative slot
  if (halted) return;
  const size_t sst = a.Stg[1] - a.Stg[0];
  const int n = a.n, S = a.tb.s, s = a.s;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p0 = blockIdx.x * kPW * ppw;              // the workgroup's first patch
  const int pb = min(p0 + wave * ppw, n), pe = min(pb + ppw, n);
  const int ne = kPW * ppw / kEW;                     // patches per epilogue wave
  const int pp = p0 + wave * ne + lane;               // this lane's epilogue patch
  const bool own = wave < kEW && lane < ne && pp < n;
  const double* X = xc ? a.Xb[1] : a.Xb[0];
  double* Xn = xc ? a.Xb[0] : a.Xb[1];
  const bool reduce = (s == 0) || (s == S - 1);
  StageOperands<NC> op;
  stage_operands<NC, true>(a, ks, sst, X, Xn, pp, own, op);
  // the rows of this wave's patches over the chunks [c0, c1) of x, copied
  // into LDS; a row's block sums are added in block order
  const int nch = lds_x_floats<TH>(a.cols) / chunk_elems<TH>();
  auto block = [&](int c0, int c1) {
    x_to_lds<TH>(static_cast<const float*>(a.vin), c0, c1, xs);
    __syncthreads();
    for (int q = pb; q < pe; ++q) {
      double acc[G];
#pragma unroll
      for (int k = 0; k < G; ++k) acc[k] = 0.0;
      gemv_lds_acc<TH, G, 2>(static_cast<const TH*>(a.K), a.ld, q * G, a.rows_all, c0, c1, xs,
                             acc);
      if constexpr (std::is_same<TK, bf16nf_t>::value) {
        const int o = a.nf_off[q];
        gemv_nf_lo_lds<G>(a.Knf_lo + (size_t)o * G * kNfChunk, a.nf_list + o,
                          a.nf_off[q + 1] - o, c0, CB ? c1 : INT_MAX, xs, acc);
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const double tot = wave_sum(acc[k]);
        if (lane == 0) red[q - p0][k] = (c0 == 0) ? tot : red[q - p0][k] + tot;
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
  __syncthreads();
  if (wave >= kEW) return;
  double r9[kNRed];
  r9_neutral(r9);
  if (own) {
    double rsum[G];
#pragma unroll
    for (int k = 0; k < G; ++k) rsum[k] = red[pp - p0][k];
    stage_epilogue<VT, NC, G, true>(a, c, kss, sst, hstep, X, Xn, pp, rsum, op, r9);
  }
  if (reduce) write_partials(a, r9, blockIdx.x * kEW + wave, gridDim.x * kEW);
}

// Additive (IMEX) step of the scalar layout, one launch per stage with the
// same wave-per-patch product as k_stage (src/rsf_imex.c with PETSc
// TSARKIMEX l2).  Slots: G_j in Stg[j], Fimpl_j in Stg[3 + j], j = 0..2;
// stage states Y_2, Y_3 in Yb[0], Yb[1].
//   s = 0  initial evaluation: G_1, Fimpl_1 at X, reductions of X
//   s = 1  G_1, Fimpl_1 at X when X is new (else kept from the rejected
//          attempt, and the product is skipped); Newton for Y_2
//   s = 2  G_2 at Y_2; Newton for Y_3 (initial guess Y_2)
//   s = 3  G_3 at Y_3; X_new = X + h sum_j b_j (G_j + Fimpl_j), the
//          embedded error, reductions of X_new and v(X_new) for the next
//          step's stage 1
// A cell whose Newton solve fails, or a stage state outside the domain,
// fails the attempt (step retried at h / 4, like PETSc's failed stage).
template <typename TK, int G>
__global__ __launch_bounds__(64 * kStagePPB, 4) void k_imex(StageArgs a) {
  using VT = typename KTraits<TK>::VT;
  constexpr int D = 4;
  __shared__ double rs[kStagePPB][G];
  DevCtl* c = a.ctl;
  const int halted = c->halted, xc = c->xcur, g1 = c->imex_g1;
  const int mon_copy = c->mon_copy, rupt_update = c->rupt_update;
  const int failed = c->stage_fail;
  const double hstep = c->h;
  const int n = a.n, s = a.s;
  // nothing to do once halted, or for the later stages of an attempt whose
  // earlier stage already failed (the controller rejects it)
  if (halted || (failed && s >= 2)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = lane;
  const int p = blockIdx.x * kStagePPB + t;
  const double* X = xc ? a.Xb[1] : a.Xb[0];
  double* Xn = xc ? a.Xb[0] : a.Xb[1];
  const bool stream = (s != 1) || g1;
  if (stream) {
    const void* vin = (s <= 1) ? (xc ? a.vxb[1] : a.vxb[0]) : a.vin;
    const int pw = min(blockIdx.x * kStagePPB + wave, n - 1);
    double rowsum[G];
    if constexpr (std::is_same<TK, bf16nf_t>::value) {
      const int o = a.nf_off[pw];
      gemv_nf_wave<G>(static_cast<const bf16_t*>(a.K), a.Knf_lo + (size_t)o * G * kNfChunk,
                      a.nf_list + o, a.nf_off[pw + 1] - o, a.ld, pw * G, a.rows_all, a.cols,
                      static_cast<const float*>(vin), rowsum);
    } else {
      gemv_wave<TK, G>(static_cast<const TK*>(a.K), a.ld, pw * G, a.rows_all, a.cols,
                       static_cast<const VT*>(vin), rowsum);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < G; ++k) rs[wave][k] = rowsum[k];
    }
    __syncthreads();
  }
  if (wave != 0) return;
  const bool reduce = (s == 0) || (s == 3);
  double r9[kNRed];
  r9[RED_ERR] = 0.0; r9[RED_VMAX] = 0.0; r9[RED_VSUM] = 0.0; r9[RED_SLIP] = 0.0;
  r9[RED_MU] = 0.0; r9[RED_SMAX] = -1e30; r9[RED_SMIN] = 1e30; r9[RED_XN2] = 0.0;
  r9[RED_DX2] = 0.0;
  if (t < kStagePPB && p < n) {
    const PatchCoef cf = load_coef(a.coef, n, p);
    const Tableau& tb = a.tb;
    double x[D], y[D];
#pragma unroll
    for (int f = 0; f < D; ++f) x[f] = X[f * n + p];
    if (s == 1) {
      // attempt start: monitor snapshot and rupture marking (as k_prep)
      if (mon_copy) {
#pragma unroll
        for (int f = 0; f < D; ++f) a.Xmon[f * n + p] = x[f];
      }
      if (rupt_update && vel_mag<1>(a.fp, cf.a, x) >= a.so.rupture_vth) a.rupt[p] = 1;
    }
    // state at which this launch evaluates the explicit part
    if (s <= 1) {
#pragma unroll
      for (int f = 0; f < D; ++f) y[f] = x[f];
    } else {
      const double* Ys = a.Yb[s - 2];
#pragma unroll
      for (int f = 0; f < D; ++f) y[f] = Ys[f * n + p];
    }
    if (stream) {
      const int j = (s <= 1) ? 0 : s - 1;
      double g[D];
      imex_explicit(a.fp, cf, y, rs[t][0], (G > 1) ? rs[t][1] : 0.0, a.sinc[p], a.sinc[n + p],
                    g);
      double* Gs = a.Stg[j];
#pragma unroll
      for (int f = 0; f < D; ++f) Gs[f * n + p] = g[f];
      if (j == 0) {
        double fi[D];
        imex_implicit(a.fp, cf, y, fi);
        double* Fs = a.Stg[3];
#pragma unroll
        for (int f = 0; f < D; ++f) Fs[f * n + p] = fi[f];
      }
    }
    if (s == 1 || s == 2) {
      // stage k = s (0-based) state: Z = X + h sum_{j<k} (At[k][j] Fimpl_j + A[k][j] G_j)
      const int k = s;
      double Z[D];
#pragma unroll
      for (int f = 0; f < D; ++f) {
        double acc = 0.0;
        for (int j = 0; j < k; ++j)
          acc += tb.At[k][j] * a.Stg[3 + j][f * n + p] + tb.A[k][j] * a.Stg[j][f * n + p];
        Z[f] = x[f] + hstep * acc;
      }
      double yk[D], fi[D];
      yk[0] = y[0];
      yk[1] = y[1];
      const int it = imex_newton(a.fp, cf, Z, hstep * tb.At[k][k], a.so.rtol, a.so.atol, yk, fi);
      if (it < 0) {
        atomicOr(&c->stage_fail, 1);
      } else {
        double* Yk = a.Yb[k - 1];
        double* Fk = a.Stg[3 + k];
#pragma unroll
        for (int f = 0; f < D; ++f) {
          Yk[f * n + p] = yk[f];
          Fk[f * n + p] = fi[f];
        }
        write_vel<VT, 1>(a.fp, cf.a, yk, static_cast<VT*>(a.vout), p, a.vpl);
        if (!rsf_patch_ok<1>(a.fp, cf, yk)) atomicOr(&c->stage_fail, 1);
      }
    } else if (s == 3) {
      // new state, embedded error and the next step's product input
      double xn[D], en = 0.0;
#pragma unroll
      for (int f = 0; f < D; ++f) {
        double acc = 0.0, ae = 0.0;
        for (int j = 0; j < 3; ++j) {
          const double gf = a.Stg[j][f * n + p] + a.Stg[3 + j][f * n + p];
          acc += tb.b[j] * gf;
          ae += tb.bh[j] * gf;
        }
        xn[f] = x[f] + hstep * acc;
        ae = x[f] + hstep * ae;
        Xn[f * n + p] = xn[f];
        const double tol = a.so.atol[f] + a.so.rtol * fmax(fabs(xn[f]), fabs(ae));
        if (tol > 0.0) en = fmax(en, fabs(xn[f] - ae) / tol);
        const double dm = a.Xmon[f * n + p] - xn[f];
        r9[RED_DX2] += dm * dm;
      }
      r9[RED_ERR] = en;
      write_vel<VT, 1>(a.fp, cf.a, xn, static_cast<VT*>(xc ? a.vxb[0] : a.vxb[1]), p, a.vpl);
#pragma unroll
      for (int f = 0; f < D; ++f) y[f] = xn[f];
    }
    if (reduce) {
      const double vm = vel_mag<1>(a.fp, cf.a, y);
      r9[RED_VMAX] = vm;
      r9[RED_VSUM] = vm;
      r9[RED_SLIP] = y[3];
      r9[RED_MU] = y[1] / y[2];
      r9[RED_SMAX] = y[2];
      r9[RED_SMIN] = y[2];
#pragma unroll
      for (int f = 0; f < D; ++f) r9[RED_XN2] += y[f] * y[f];
    }
  }
  if (reduce) {
#pragma unroll
    for (int k = 0; k < kNRed; ++k) {
      const double q = wave_reduce(r9[k], red_kind(k));
      if (t == 0) a.partials[(size_t)k * gridDim.x + blockIdx.x] = q;
    }
  }
}

template <int NC>
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl(StageArgs a, int nb) {
  ctrl_block<NC>(a, a.s, nb);
}

// ------------------------------------------------------------ host side

void fill_tableau(Tableau& tb, int method) {
  memset(&tb, 0, sizeof(tb));
  if (method == RSFG_ARKIMEX_L2) {
    // PETSc 3.19.6 TSARKIMEXL2 (src/ts/impls/arkimex/arkimex.c; the
    // reference's -imex default, src/rsf_imex.c:283): explicit first stage,
    // L-stable implicit part with gamma = 1 - 1/sqrt 2, b = bt = last
    // implicit row, first-order embedded bembedt, quadratic dense output
    const double s2 = sqrt(2.0), g = 1.0 - 1.0 / s2;
    tb.s = 3;
    tb.order = 2;
    tb.imex = 1;
    tb.A[1][0] = 2.0 - s2;
    tb.A[2][0] = 0.5; tb.A[2][1] = 0.5;
    tb.At[1][0] = g; tb.At[1][1] = g;
    tb.At[2][0] = 1.0 / (2.0 * s2); tb.At[2][1] = 1.0 / (2.0 * s2); tb.At[2][2] = g;
    for (int j = 0; j < 3; ++j) tb.b[j] = tb.At[2][j];
    tb.bh[0] = (4.0 - s2) / 8.0; tb.bh[1] = (4.0 - s2) / 8.0; tb.bh[2] = 1.0 / (2.0 * s2);
    tb.d1[0] = 1.0 / s2; tb.d1[1] = 1.0 / s2; tb.d1[2] = 1.0 - s2;
    tb.d2[0] = -1.0 / (2.0 * s2); tb.d2[1] = -1.0 / (2.0 * s2); tb.d2[2] = 1.0 / s2;
  } else if (method == RSFG_RK3BS) {
    tb.s = 4;
    tb.order = 3;
    tb.A[1][0] = 0.5;
    tb.A[2][1] = 0.75;
    tb.A[3][0] = 2.0 / 9.0; tb.A[3][1] = 1.0 / 3.0; tb.A[3][2] = 4.0 / 9.0;
    tb.b[0] = 2.0 / 9.0; tb.b[1] = 1.0 / 3.0; tb.b[2] = 4.0 / 9.0;
    tb.bh[0] = 7.0 / 24.0; tb.bh[1] = 0.25; tb.bh[2] = 1.0 / 3.0; tb.bh[3] = 0.125;
  } else {
    tb.s = 7;
    tb.order = 5;
    tb.A[1][0] = 0.2;
    tb.A[2][0] = 3.0 / 40.0; tb.A[2][1] = 9.0 / 40.0;
    tb.A[3][0] = 44.0 / 45.0; tb.A[3][1] = -56.0 / 15.0; tb.A[3][2] = 32.0 / 9.0;
    tb.A[4][0] = 19372.0 / 6561.0; tb.A[4][1] = -25360.0 / 2187.0;
    tb.A[4][2] = 64448.0 / 6561.0; tb.A[4][3] = -212.0 / 729.0;
    tb.A[5][0] = 9017.0 / 3168.0; tb.A[5][1] = -355.0 / 33.0; tb.A[5][2] = 46732.0 / 5247.0;
    tb.A[5][3] = 49.0 / 176.0; tb.A[5][4] = -5103.0 / 18656.0;
    tb.A[6][0] = 35.0 / 384.0; tb.A[6][2] = 500.0 / 1113.0; tb.A[6][3] = 125.0 / 192.0;
    tb.A[6][4] = -2187.0 / 6784.0; tb.A[6][5] = 11.0 / 84.0;
    for (int j = 0; j < 7; ++j) tb.b[j] = tb.A[6][j];
    tb.bh[0] = 5179.0 / 57600.0; tb.bh[2] = 7571.0 / 16695.0; tb.bh[3] = 393.0 / 640.0;
    tb.bh[4] = -92097.0 / 339200.0; tb.bh[5] = 187.0 / 2100.0; tb.bh[6] = 1.0 / 40.0;
  }
}

struct Launcher {
  virtual ~Launcher() {}
  virtual void init_vel(const StageArgs& a, hipStream_t st) = 0;
  virtual void prep(const StageArgs& a, hipStream_t st) = 0;
  virtual void stage(const StageArgs& a, hipStream_t st) = 0;
  virtual void ctrl(const StageArgs& a, hipStream_t st) = 0;
  virtual void imex(const StageArgs& a, hipStream_t st) = 0;
  // columns of partials[kNRed][.] the reducing stages write (k_ctrl reads
  // that many)
  virtual int red_cols() const = 0;
};

// Persistent stage kernel geometry for n patches and cols columns: ppw
// patches per wave and the workgroup count, or nwg = 0 where the per-patch
// kernel runs: FP64 entries, more than kMaxPPW patches per wave, fewer than
// kMinPPW (the persistent kernel's fixed cost, about 11 us per launch of x
// copy, operand requests and the epilogue after the last stream, then
// outweighs its faster stream: measured 33 against 31 us at 2 patches per
// wave, 318 against 361 us at 8).  x longer than the LDS copy runs in
// column blocks.  RSFG_STAGE_PERSIST=0 in the environment selects the
// per-patch kernel everywhere, =1 the persistent kernel wherever it can run.
constexpr int kMinPPW = 3;

template <typename TK>
static void persist_geometry(int ncu, int n, int cols, bool allow, int* nwg, int* ppw) {
  *nwg = 0;
  *ppw = 0;
  if constexpr (std::is_same<typename KTraits<TK>::VT, float>::value) {
    const char* e = getenv("RSFG_STAGE_PERSIST");
    if (!allow || (e && e[0] == '0')) return;
    (void)cols;   // x longer than the LDS copy runs in column blocks
    const bool force = e && e[0] == '1';
    if (!persist_split(ncu, n, kMaxPPW, nwg, ppw) || (!force && *ppw < kMinPPW)) *nwg = *ppw = 0;
  }
}

template <typename TK, int NC, int G>
struct LauncherT : Launcher {
  using VT = typename KTraits<TK>::VT;
  using TH = typename std::conditional<std::is_same<TK, bf16nf_t>::value, bf16_t, TK>::type;
  int nwg_p = 0, ppw_p = 0, n_ = 0;
  bool blocks_ = false;   // x longer than the LDS copy: column blocks
  LauncherT(int ncu, int n, int cols, bool allow_persist) : n_(n) {
    persist_geometry<TK>(ncu, n, cols, allow_persist, &nwg_p, &ppw_p);
    blocks_ = lds_blocks<TH>(cols) > 1;
  }
  void init_vel(const StageArgs& a, hipStream_t st) override {
    hipLaunchKernelGGL((k_init_vel<VT, NC>), dim3((a.n + 255) / 256), dim3(256), 0, st, a);
  }
  void prep(const StageArgs& a, hipStream_t st) override {
    hipLaunchKernelGGL((k_prep<VT, NC>), dim3((a.n + 255) / 256), dim3(256), 0, st, a);
  }
  void stage(const StageArgs& a, hipStream_t st) override {
    if constexpr (std::is_same<VT, float>::value) {
      if (nwg_p > 0) {
        if (blocks_)
          hipLaunchKernelGGL((k_stage_p<TK, NC, G, true>), dim3(nwg_p), dim3(64 * kPW), 0, st,
                             a, ppw_p);
        else
          hipLaunchKernelGGL((k_stage_p<TK, NC, G, false>), dim3(nwg_p), dim3(64 * kPW), 0, st,
                             a, ppw_p);
        return;
      }
    }
    hipLaunchKernelGGL((k_stage<TK, NC, G>), dim3(wg_per_patch(a.n)), dim3(64 * kStagePPB), 0,
                       st, a);
  }
  void ctrl(const StageArgs& a, hipStream_t st) override {
    hipLaunchKernelGGL((k_ctrl<NC>), dim3(1), dim3(kCtrlThreads), 0, st, a, red_cols());
  }
  // additive stages exist for the scalar layout only (the solve rejects
  // the method otherwise); their launcher is built without the persistent
  // kernel, so k_ctrl reads the per-patch kernel's columns
  void imex(const StageArgs& a, hipStream_t st) override {
    if constexpr (NC == 1) {
      hipLaunchKernelGGL((k_imex<TK, G>), dim3(wg_per_patch(a.n)), dim3(64 * kStagePPB), 0, st,
                         a);
    }
  }
  // one patch per wave, kStagePPB waves per workgroup
  static int wg_per_patch(int n) { return (n + kStagePPB - 1) / kStagePPB; }
  int red_cols() const override { return nwg_p > 0 ? nwg_p * kEW : wg_per_patch(n_); }
};

template <typename TK>
static Launcher* make_launcher_t(int ncomp, int G, int ncu, int n, int cols, bool allow) {
  if (ncomp == 1) {
    if (G == 1) return new LauncherT<TK, 1, 1>(ncu, n, cols, allow);
    return new LauncherT<TK, 1, 2>(ncu, n, cols, allow);
  }
  if (G == 2) return new LauncherT<TK, 2, 2>(ncu, n, cols, allow);
  return new LauncherT<TK, 2, 3>(ncu, n, cols, allow);
}

// allow_persist: false for the additive (IMEX) method, whose stage kernel
// is per-patch
static Launcher* make_launcher(const rsfg_ctx* h, int prec, bool allow_persist) {
  const int nc = h->ncomp, G = h->G, n = h->n, cols = h->cols, cu = h->ncu;
  switch (prec) {
    case RSFG_PREC_FP32: return make_launcher_t<float>(nc, G, cu, n, cols, allow_persist);
    case RSFG_PREC_BF16X2: return make_launcher_t<bf16x2_t>(nc, G, cu, n, cols, allow_persist);
    case RSFG_PREC_BF16: return make_launcher_t<bf16_t>(nc, G, cu, n, cols, allow_persist);
    case RSFG_PREC_BF16NF: return make_launcher_t<bf16nf_t>(nc, G, cu, n, cols, allow_persist);
    default: return make_launcher_t<double>(nc, G, cu, n, cols, allow_persist);
  }
}

void fill_stage_args(rsfg_ctx* h, const rsfg_solver_opts* o, const void* K, StageArgs& a) {
  const int n = h->n;
  memset(&a, 0, sizeof(a));
  a.K = K;
  a.Knf_lo = h->Knf_lo;
  a.nf_off = h->nf_off;
  a.nf_list = h->nf_list;
  a.ld = h->ld;
  a.rows_all = h->rows_all;
  a.cols = h->cols;
  a.n = n;
  a.G = h->G;
  a.coef = h->d_coef;
  // d_sinc: [full shear | full normal | additive shear | additive normal | vpl]
  const size_t nl = (size_t)(h->ncomp + 1) * n;
  const bool fold = (o->prec != RSFG_PREC_FP64);
  a.sinc = fold ? h->d_sinc + nl : h->d_sinc;
  a.vpl = fold ? h->d_sinc + 2 * nl : nullptr;
  a.fp = h->fp;
  for (int i = 0; i < 2; ++i) {
    a.Xb[i] = h->Xb[i];
    a.Yb[i] = h->Yb[i];
    a.Pb[i] = h->Pb[i];
  }
  for (int i = 0; i < 8; ++i) a.Stg[i] = h->Stg[i];
  a.vxb[0] = h->vxb[0];
  a.vxb[1] = h->vxb[1];
  a.Xmon = h->Xmon;
  a.partials = h->partials;
  a.ev = h->d_ev;
  a.mon = h->d_mon;
  a.tau0 = h->d_cat;
  a.slip0 = h->d_cat + n;
  a.rupt = h->d_rupt;
  a.catr = h->d_catr;
  a.area = h->d_area;
  a.ctl = h->d_ctl;
  a.hlog = h->d_hlog;
  a.hlog_cap = h->hlog_cap;
  fill_tableau(a.tb, o->method);
  a.so.rtol = o->rtol;
  for (int f = 0; f < 6; ++f) a.so.atol[f] = o->atol[f];
  a.so.dt_monitor = o->dt_monitor;
  a.so.rdx = o->rdx_monitor;
  a.so.adx = o->adx_monitor;
  a.so.vel_event = o->vel_event;
  a.so.hyst = o->vel_event_hyst;
  a.so.rupture_vth = o->rupture_vth;
  a.so.shear_mu = h->scale * 1e4;   // scale = G_SI / 1e4 (src/rsf_solve.c:235-237)
  a.so.max_domain_reject = o->max_domain_reject;
}

// controller state for a solve from the handle's current (t, X); then K_0
static int start_solve(rsfg_ctx* h, const rsfg_solver_opts* o, double t_end, StageArgs& a,
                       Launcher* L) {
  const int n = h->n;
  DevCtl* c = h->h_ctl;
  memset(c, 0, sizeof(*c));
  c->t = h->t;
  c->h = o->dt_init;
  c->t_end = t_end;
  c->hmax = (o->dt_max < o->dt_monitor) ? o->dt_max : o->dt_monitor;
  c->max_steps = o->max_steps;
  for (int i = 0; i < 8; ++i) c->kslot[i] = i;
  c->max_event = o->max_events > 0 ? o->max_events : kMaxEv;
  c->max_mon = o->max_monitor > 0 ? o->max_monitor : kMaxMon;
  c->xcur = h->xcur;
  RSFG_HIP(hipMemcpyAsync(h->d_ctl, c, sizeof(DevCtl), hipMemcpyHostToDevice, h->stream));
  // the GEMV reads the input vector up to a whole chunk past the last
  // column: its padding must be zero in this precision's element type
  // (vin[0], vin[1], vxb[0], vxb[1] are consecutive)
  RSFG_HIP(hipMemsetAsync(h->vin[0], 0, 4 * h->vin_elems * sizeof(double), h->stream));
  RSFG_HIP(hipMemsetAsync(h->d_rupt, 0, sizeof(int) * n, h->stream));
  a.s = 0;
  if (a.tb.imex) {
    // G_1 and Fimpl_1 at the initial state
    a.vout = h->vxb[h->xcur];
    L->init_vel(a, h->stream);
    a.vout = nullptr;
    L->imex(a, h->stream);
  } else {
    // K_0 at the initial state
    a.vout = h->vin[0];
    L->init_vel(a, h->stream);
    a.vin = h->vin[0];
    a.vout = nullptr;
    L->stage(a, h->stream);
  }
  L->ctrl(a, h->stream);
  RSFG_HIP(hipGetLastError());
  RSFG_HIP(hipMemcpyAsync(c, h->d_ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  if (c->status == RSFG_EDOMAIN) {
    set_error("rsfg_solve: initial state fails the domain check");
    return RSFG_EDOMAIN;
  }
  return RSFG_OK;
}

void drop_graph(rsfg_ctx* h) {
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->ggraph) hipGraphDestroy(h->ggraph);
  h->gexec = nullptr;
  h->ggraph = nullptr;
  h->gkey.clear();
}

// Captures one batch of step attempts into h's cached graph unless the
// cached one was captured from the same argument block and batch size.  On
// a failure the capture is ended and nothing is cached.
template <typename F>
static int cached_graph(rsfg_ctx* h, const StageArgs& a, int per, F enqueue) {
  std::vector<unsigned char> key(sizeof(StageArgs) + sizeof(int));
  memcpy(key.data(), &a, sizeof(StageArgs));
  memcpy(key.data() + sizeof(StageArgs), &per, sizeof(int));
  if (h->gexec && key == h->gkey) return RSFG_OK;
  drop_graph(h);
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "rsfg_solve: begin capture", __FILE__, __LINE__);
  enqueue();
  hipGraph_t g = nullptr;
  // end the capture whatever the launches returned, so the stream leaves
  // capture mode
  const hipError_t el = hipGetLastError();
  e = hipStreamEndCapture(h->stream, &g);
  if (e == hipSuccess) e = el;
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(e, "rsfg_solve: capture of the attempt graph", __FILE__, __LINE__);
  }
  h->ggraph = g;
  h->gexec = x;
  h->gkey.swap(key);
  return RSFG_OK;
}

// cont = false: a solve from the handle's current (t, X); cont = true: the
// continuation of the last solve (rsfg_solve_continue), which keeps the
// controller block, the FSAL derivative, the monitor reference and the
// event state and appends to the logs
int solve_device(rsfg_ctx* h, const rsfg_solver_opts* o, double t_end, rsfg_stats* st,
                 bool cont) {
  const auto wall0 = std::chrono::steady_clock::now();
  const void* K = operator_ptr(h, o->prec);
  RSFG_CHECK(K != nullptr, RSFG_ESTATE, "rsfg_solve: operator precision not assembled");
  RSFG_CHECK(o->method == RSFG_RK5DP || o->method == RSFG_RK3BS ||
                 o->method == RSFG_ARKIMEX_L2,
             RSFG_EINVAL, "rsfg_solve: unknown method");
  RSFG_CHECK(o->method != RSFG_ARKIMEX_L2 || h->ncomp == 1, RSFG_EINVAL,
             "rsfg_solve: the IMEX method is defined for the scalar slip layout "
             "(RSFG_SLIP_STRIKE / DIP / RAKE), as in the reference");
  RSFG_CHECK(o->max_events <= (1 << 24) && o->max_monitor <= (1 << 26), RSFG_EINVAL,
             "rsfg_solve: max_events at most 2^24 and max_monitor at most 2^26 rows");
  if (cont) {
    RSFG_CHECK(h->cont_ok, RSFG_ESTATE,
               "rsfg_solve_continue: no solve to continue (none yet, or the state, "
               "friction, loading or solver buffers changed since)");
    RSFG_CHECK(o->method == h->cont_method && o->prec == h->cont_prec, RSFG_EINVAL,
               "rsfg_solve_continue: method and precision must be those of the solve "
               "being continued");
  }
  h->cont_ok = false;
  const int ev_rows = o->max_events > 0 ? o->max_events : kMaxEv;
  const int mon_rows = o->max_monitor > 0 ? o->max_monitor : kMaxMon;
  {
    // a continuation keeps the rows already logged (ensure_logs copies them)
    const int rc0 = ensure_logs(h, ev_rows, mon_rows);
    if (rc0 != RSFG_OK) return rc0;
  }
  Launcher* L = make_launcher(h, o->prec, o->method != RSFG_ARKIMEX_L2);
  if ((size_t)L->red_cols() * kNRed > h->partial_count) {
    delete L;
    set_error("rsfg_solve: partial buffer too small");
    return RSFG_ESTATE;
  }
  StageArgs a;
  fill_stage_args(h, o, K, a);
  DevCtl* c = h->h_ctl;
  if (cont) {
    // h_ctl holds the block as the last batch left it (the halted solve
    // changes nothing after halting)
    RSFG_CHECK(c->max_event <= ev_rows && c->max_mon <= mon_rows, RSFG_EINVAL,
               "rsfg_solve_continue: the logs cannot shrink below the continued solve's");
    c->t_end = t_end;
    c->hmax = (o->dt_max < o->dt_monitor) ? o->dt_max : o->dt_monitor;
    c->max_event = ev_rows;
    c->max_mon = mon_rows;
    c->max_steps = (o->max_steps > LLONG_MAX - c->steps) ? LLONG_MAX : c->steps + o->max_steps;
    c->halted = (c->t < t_end) ? 0 : 1;
    c->status = 0;
    const hipError_t e =
        hipMemcpyAsync(h->d_ctl, c, sizeof(DevCtl), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) {
      delete L;
      return hip_fail(e, "rsfg_solve_continue: control block", __FILE__, __LINE__);
    }
  } else {
    const int rc0 = start_solve(h, o, t_end, a, L);
    if (rc0 != RSFG_OK) {
      delete L;
      return rc0;
    }
  }
  // one batch = `per` step attempts; captured once into a graph and
  // replayed (RSFG_NO_GRAPH=1 launches the same kernels directly, e.g.
  // under tools that do not follow graph launches)
  const int S = a.tb.s;
  const int per = o->steps_per_launch > 0 ? o->steps_per_launch : 32;
  const char* ng = getenv("RSFG_NO_GRAPH");
  const bool use_graph = !(ng && ng[0] == '1');
  auto enqueue_batch = [&]() {
    for (int k = 0; k < per && a.tb.imex; ++k) {
      StageArgs b = a;
      b.s = 1;
      b.vout = h->vin[0];
      L->imex(b, h->stream);
      b.s = 2;
      b.vin = h->vin[0];
      b.vout = h->vin[1];
      L->imex(b, h->stream);
      b.s = 3;
      b.vin = h->vin[1];
      b.vout = nullptr;
      L->imex(b, h->stream);
      L->ctrl(b, h->stream);
    }
    for (int k = 0; k < per && !a.tb.imex; ++k) {
      StageArgs b = a;
      b.s = 0;
      b.vout = h->vin[1];
      L->prep(b, h->stream);
      for (int s = 1; s < S; ++s) {
        b.s = s;
        b.vin = h->vin[s & 1];
        b.vout = (s < S - 1) ? h->vin[(s + 1) & 1] : nullptr;
        L->stage(b, h->stream);
      }
      L->ctrl(b, h->stream);
    }
  };
  int rc = RSFG_OK;
  if (use_graph && !c->halted) rc = cached_graph(h, a, per, enqueue_batch);
  long launches = 0;
  while (rc == RSFG_OK && !c->halted) {
    hipError_t e = hipSuccess;
    if (use_graph) e = hipGraphLaunch(h->gexec, h->stream);
    else {
      enqueue_batch();
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c, h->d_ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      rc = hip_fail(e, "rsfg_solve batch launch", __FILE__, __LINE__);
      break;
    }
    ++launches;
  }
  delete L;
  if (rc != RSFG_OK) return rc;
  h->xcur = c->xcur;
  h->t = c->t;
  h->nevent = c->nevent;
  h->nmon = c->nmon;
  h->ncat = c->ncat;
  if (st) {
    memset(st, 0, sizeof(*st));
    st->steps = c->steps;
    st->rejects = c->rejects;
    st->domain_rejects = c->domain_rejects;
    st->rhs_evals = c->rhs_evals;
    st->launches = launches;
    st->t = c->t;
    st->dt = c->h;
    st->status = c->status == 2 ? RSFG_OK : c->status;
    st->nevent = c->nevent;
    st->nmonitor = c->nmon;
    st->events_dropped = c->ev_drop;
    st->monitor_dropped = c->mon_drop;
    st->ncatalog = c->ncat;
    st->catalog_dropped = c->cat_drop;
    st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
  }
  if (c->status == RSFG_ESTORM) {
    set_error("rsfg_solve: rejection storm (consecutive out-of-domain trial steps)");
    return RSFG_ESTORM;
  }
  if (c->status == RSFG_EDOMAIN) {
    set_error("rsfg_solve: step size underflow or non-finite time");
    return RSFG_EDOMAIN;
  }
  // a solve that ended at t_end or at max_steps can be continued
  h->cont_ok = true;
  h->cont_method = o->method;
  h->cont_prec = o->prec;
  return RSFG_OK;
}

// Mean duration of the fused stage kernel (stage 1 of rk5dp: K v(Y_1), the
// RHS epilogue and Y_2) at the current state.  Repeated launches read the
// same inputs and rewrite the same outputs, so every launch does identical
// work; the state of the handle is left as it was.
int stage_bench(rsfg_ctx* h, int prec, int nrep, double* ms) {
  const void* K = operator_ptr(h, prec);
  RSFG_CHECK(K != nullptr, RSFG_ESTATE, "rsfg_stage_bench: precision not assembled");
  rsfg_solver_opts o;
  rsfg_solver_defaults(&o);
  o.prec = prec;
  h->cont_ok = false;
  Launcher* L = make_launcher(h, prec, true);
  StageArgs a;
  fill_stage_args(h, &o, K, a);
  int rc = start_solve(h, &o, h->t + 1e30, a, L);
  if (rc == RSFG_OK) {
    StageArgs b = a;
    b.s = 0;
    b.vout = h->vin[1];
    L->prep(b, h->stream);
    b.s = 1;
    b.vin = h->vin[1];
    b.vout = h->vin[0];
    L->stage(b, h->stream);  // warm
    hipEventRecord(h->ev0, h->stream);
    for (int i = 0; i < nrep; ++i) L->stage(b, h->stream);
    hipEventRecord(h->ev1, h->stream);
    hipError_t e = hipEventSynchronize(h->ev1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, h->ev0, h->ev1);
    if (e != hipSuccess) rc = hip_fail(e, "rsfg_stage_bench", __FILE__, __LINE__);
    *ms = t / nrep;
  }
  delete L;
  return rc;
}

}  // namespace rsfg
