// Device pieces shared by the single-fault solver and the ensemble: the
// stage argument block, per-patch helpers, block reductions, the dense
// output and the step controller, and the bodies of the per-attempt
// preparation and control kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gemv_dev.h"
#include "rsf_dev.h"
#include "rsfg_common.h"

namespace rsfg {

constexpr double kSecPerYear = 31557600.0;

struct StageArgs {
  const void* K;
  const bf16_t* Knf_lo;      // bf16 + near-field lo format: lo blocks,
  const int* nf_off;         //   per-patch chunk offsets (n + 1)
  const uint16_t* nf_list;   //   and chunk lists
  long ld;
  int rows_all, cols, n, G;
  const double* coef;  // a, b, dc
  const double* sinc;  // loading: ncomp n shear rows, then n normal
  const double* vpl;   // plate rate per patch when folded into K v, else null
  FricParams fp;
  double* Xb[2];
  double* Yb[2];
  double* Pb[2];       // state-only RHS records of the stage states
                       // (PreDim<NC>::v x n each; stage s reads Pb[s & 1])
  double* Stg[8];
  double* Xmon;
  const void* vin;
  void* vout;
  void* vxb[2];        // additive pair: product input v(X) of each X buffer
  double* partials;
  double* ev;
  double* mon;
  double* tau0;
  double* slip0;
  int* rupt;
  double* catr;        // catalog rows (kCatCols each, capacity max_event)
  const double* area;  // per-patch catalog area 4 l w
  DevCtl* ctl;
  double* hlog;        // accepted step sizes, in order (optional; the control
  long long hlog_cap;  // block the host copies holds no device address)
  Tableau tb;
  SolveOpts so;
  int s;       // stage index; 0 = evaluation of K_0 at the initial state
};

template <int NC>
struct Dim {
  static constexpr int v = 2 * NC + 2;
};

__device__ __forceinline__ PatchCoef load_coef(const double* coef, int n, int p) {
  PatchCoef c;
  c.a = coef[p];
  c.b = coef[n + p];
  c.dc = coef[2 * n + p];
  return c;
}

template <int NC>
__device__ __forceinline__ double vel_mag(const FricParams& fp, double a, const double* s) {
  double v[2];
  rsf_patch_vel<NC>(fp, a, s, v);
  return (NC == 1) ? fabs(v[0]) : sqrt(v[0] * v[0] + v[1] * v[1]);
}

// product input of patch p: its slip rate, minus the plate rate along the
// slip direction when the backslip loading is folded into the product
// (vpl != nullptr, reduced-precision operators: K (v - vpl) instead of
// K v + K (-vpl), so the storage rounding scales with the deviation from
// plate rate instead of acting as a spurious loading)
template <typename VT, int NC>
__device__ __forceinline__ void write_vel(const FricParams& fp, double a,
                                          const double* s, VT* vout, int p,
                                          const double* vpl) {
  double v[2];
  rsf_patch_vel<NC>(fp, a, s, v);
  if (vpl) v[0] -= vpl[p];
#pragma unroll
  for (int c = 0; c < NC; ++c) vout[NC * p + c] = static_cast<VT>(v[c]);
}

// the same product input from a stage state's state-only record
template <typename VT, int NC>
__device__ __forceinline__ void write_vel_pre(const double* P, VT* vout, int p,
                                              const double* vpl) {
  double v[2] = {P[0], NC == 2 ? P[1] : 0.0};
  if (vpl) v[0] -= vpl[p];
#pragma unroll
  for (int c = 0; c < NC; ++c) vout[NC * p + c] = static_cast<VT>(v[c]);
}

template <int NC>
__device__ __forceinline__ void store_pre(const double* P, double* Pb, int n, int p) {
#pragma unroll
  for (int k = 0; k < PreDim<NC>::v; ++k) Pb[k * n + p] = P[k];
}

// ------------------------------------------------------------ block helpers
__device__ __forceinline__ double red_op(int op, double v, double o) {
  return (op == 0) ? v + o : (op == 1 ? fmax(v, o) : fmin(v, o));
}

// combination of reduction slot k (RED_*): 0 sum, 1 max, 2 min
__device__ __forceinline__ int red_kind(int k) {
  return (k == RED_ERR || k == RED_VMAX || k == RED_SMAX) ? 1 : (k == RED_SMIN ? 2 : 0);
}

__device__ __forceinline__ double red_identity(int op) {
  return op == 0 ? 0.0 : (op == 1 ? -1e30 : 1e30);
}

// op 0 sum, 1 max, 2 min over the whole workgroup (up to 16 waves, sh holds
// 16 doubles); fixed combination order, so sums are reproducible
__device__ __forceinline__ double block_reduce(double v, int op, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red_op(op, v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double r = sh[0];
  for (int w = 1; w < nw; ++w) r = red_op(op, r, sh[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double evfun(double vmax, double thr) {
  return log10(vmax > 1e-300 ? vmax : 1e-300) - log10(thr);
}

// dense output of patch p at fraction th of the step [t, t + hs]
template <int NC>
__device__ void dense_patch(const StageArgs& a, const double* X0, const double* X1,
                            const double* const* Kp, double hs, double th, int p,
                            double* out) {
  constexpr int D = Dim<NC>::v;
  const int n = a.n;
  if (a.tb.imex) {
    // additive pair: X0 + hs sum_j (d1_j th + d2_j th^2) (G_j + Fimpl_j),
    // G_j in Kp[j], Fimpl_j in stage slot 3 + j
#pragma unroll
    for (int f = 0; f < D; ++f) {
      const int i = f * n + p;
      double acc = 0.0;
      for (int j = 0; j < 3; ++j)
        acc += (a.tb.d1[j] * th + a.tb.d2[j] * th * th) * (Kp[j][i] + a.Stg[3 + j][i]);
      out[f] = X0[i] + hs * acc;
    }
    return;
  }
  if (a.tb.s == 7) {
    const double d1 = -12715105075.0 / 11282082432.0, d3 = 87487479700.0 / 32700410799.0,
                 d4 = -10690763975.0 / 1880347072.0, d5 = 701980252875.0 / 199316789632.0,
                 d6 = -1453857185.0 / 822651844.0, d7 = 69997945.0 / 29380423.0;
#pragma unroll
    for (int f = 0; f < D; ++f) {
      const int i = f * n + p;
      const double r2 = X1[i] - X0[i];
      const double r3 = hs * Kp[0][i] - r2;
      const double r4 = r2 - hs * Kp[6][i] - r3;
      const double r5 = hs * (d1 * Kp[0][i] + d3 * Kp[2][i] + d4 * Kp[3][i] +
                              d5 * Kp[4][i] + d6 * Kp[5][i] + d7 * Kp[6][i]);
      out[f] = X0[i] + th * (r2 + (1.0 - th) * (r3 + th * (r4 + (1.0 - th) * r5)));
    }
  } else {
    const double t2 = th * th, t3 = t2 * th;
#pragma unroll
    for (int f = 0; f < D; ++f) {
      const int i = f * n + p;
      out[f] = (2 * t3 - 3 * t2 + 1) * X0[i] + (t3 - 2 * t2 + th) * hs * Kp[0][i] +
               (-2 * t3 + 3 * t2) * X1[i] + (t3 - t2) * hs * Kp[3][i];
    }
  }
}

// block-wide max |v| at dense-output fraction th
template <int NC>
__device__ double dense_vmax(const StageArgs& a, const double* X0, const double* X1,
                             const double* const* Kp, double hs, double th, double* sh) {
  constexpr int D = Dim<NC>::v;
  double m = 0.0, y[D];
  for (int p = threadIdx.x; p < a.n; p += blockDim.x) {
    dense_patch<NC>(a, X0, X1, Kp, hs, th, p, y);
    m = fmax(m, vel_mag<NC>(a.fp, a.coef[p], y));
  }
  return block_reduce(m, 1, sh);
}

// ------------------------------------------------------------ controller
// Runs in the single workgroup of k_ctrl after the final stage (or after the
// initial K_0 evaluation when s == 0) on an LDS copy c of the control
// block.  red[] holds the grid reductions.
template <int NC>
__device__ void controller(const StageArgs& a, int s, const double* red, DevCtl* c) {
  constexpr int D = Dim<NC>::v;
  __shared__ double sh[16];
  __shared__ double s_par[8];
  __shared__ int s_int[4];
  const int n = a.n, S = a.tb.s, tid = threadIdx.x;
  const SolveOpts& o = a.so;
  if (s == 0) {
    if (tid == 0) {
      if (c->stage_fail) {
        c->status = RSFG_EDOMAIN;
        c->halted = 1;
      }
      c->rhs_evals += 1;
      c->slipping = red[RED_VMAX] > o.vel_event;
      c->vprev = red[RED_VMAX];
      c->peakv = 0.0;
      c->last_vmax = red[RED_VMAX];
      // first monitor row (old time forced to t0 - 2 dt_monitor)
      if (c->nmon < c->max_mon) {
        double* row = a.mon + 10 * c->nmon;
        row[0] = (double)c->steps; row[1] = c->t; row[2] = c->t / kSecPerYear;
        row[3] = c->h; row[4] = log10(red[RED_VMAX]); row[5] = red[RED_SLIP] / n;
        row[6] = red[RED_MU] / n; row[7] = red[RED_SMAX]; row[8] = red[RED_SMIN];
        const double mv = red[RED_VSUM] / n;
        row[9] = log10(mv > 1e-300 ? mv : 1e-300);
        c->nmon++;
      } else {
        c->mon_drop++;
      }
      c->t_mon = c->t;
      c->mon_copy = 1;
      c->rupt_update = 0;
      c->stage_fail = 0;
      c->attempt_failed = 0;
      c->ev_open = 0;
      if (!(c->t < c->t_end)) c->halted = 1;
    }
    return;
  }
  // ---- accept / reject (thread 0), PETSc TSAdapt basic
  if (tid == 0) {
    // product evaluations of the attempt (additive pair: stages 2 and 3,
    // plus stage 1 after an accepted step)
    c->rhs_evals += a.tb.imex ? 2 + c->imex_g1 : S - 1;
    int acc;
    double hnew;
    if (c->stage_fail) {
      c->domain_rejects++;
      c->rejects++;
      hnew = c->h * 0.25;
      c->attempt_failed = 1;
      acc = 0;
    } else {
      double en = red[RED_ERR];
      if (!(en == en) || isinf(en)) en = 1e300;
      double safety = 0.9;
      acc = 1;
      if (en > 1.0) {
        if (c->attempt_failed) safety *= 0.5;
        acc = 0;
      }
      double hfac = (en > 0.0) ? safety * pow(en, -1.0 / a.tb.order) : 2.0;
      hfac = fmin(fmax(hfac, 0.5), 2.0);
      hnew = fmin(c->h * hfac, c->hmax);
      if (!acc) {
        c->rejects++;
        c->attempt_failed = 1;
      }
    }
    // rejection-storm guard: consecutive attempts that fail the domain
    // check (-domain_check_max_reject); step-size underflow ends the solve
    if (c->stage_fail) c->storm++;
    else c->storm = 0;
    if (!acc) {
      if (o.max_domain_reject > 0 && c->storm >= o.max_domain_reject) {
        c->status = RSFG_ESTORM;
        c->halted = 1;
      }
      if (!(hnew > 0.0) || !(c->t + hnew != c->t) || !isfinite(c->t)) {
        c->status = RSFG_EDOMAIN;
        c->halted = 1;
      }
    }
    s_int[0] = acc;
    s_par[0] = hnew;
    c->imex_g1 = 0;
    c->mon_copy = 0;
    c->rupt_update = 0;
    c->stage_fail = 0;
    if (!acc) c->h = hnew;
  }
  __syncthreads();
  if (!s_int[0]) return;
  // ---- accepted: events on the dense output of [t, t + h]
  const int xo = c->xcur;
  const double* X0 = a.Xb[xo];
  const double* X1 = a.Xb[xo ^ 1];
  const double* Kp[kMaxStages];
  for (int j = 0; j < S; ++j) Kp[j] = a.Stg[c->kslot[j]];
  const double hs = c->h;
  const double vnew = red[RED_VMAX];
  double th_lo = 0.0;
  for (int guard = 0; guard < 8; ++guard) {
    // thread 0 decides whether [th_lo, 1] holds a crossing of the event
    // function f = log10(max|v|) - log10(thr).  Both ends clearly on one side
    // of thr (relative margin 1e-9, far above the rounding of the logs)
    // cannot change sign; only otherwise are the logs evaluated.
    if (tid == 0) {
      const double thr = o.vel_event * (c->slipping ? o.hyst : 1.0);
      const double vp = c->vprev;
      const double lo_t = thr * (1.0 - 1e-9), hi_t = thr * (1.0 + 1e-9);
      bool cross = false;
      if (!((vp < lo_t && vnew < lo_t) || (vp > hi_t && vnew > hi_t))) {
        const double fnew = evfun(vnew, thr);
        const double fp = evfun(vp, thr);
        cross = (fp < 0.0 && fnew > 0.0) || (fp > 0.0 && fnew < 0.0) || fnew == 0.0;
        s_par[1] = th_lo; s_par[2] = fp; s_par[3] = 1.0; s_par[4] = fnew; s_par[5] = thr;
      }
      s_int[1] = cross;
      if (!cross) c->vprev = vnew;
    }
    __syncthreads();
    if (!s_int[1]) break;
    // Illinois regula falsi on the fraction
    int side = 0;
    for (int it = 0; it < 200; ++it) {
      if (tid == 0) {
        const double lo = s_par[1], flo = s_par[2], hi = s_par[3], fhi = s_par[4];
        double cth = (fhi - flo != 0.0) ? hi - fhi * (hi - lo) / (fhi - flo) : 0.5 * (lo + hi);
        if (!(cth > lo && cth < hi)) cth = 0.5 * (lo + hi);
        s_par[6] = cth;
      }
      __syncthreads();
      const double cth = s_par[6];
      const double vm = dense_vmax<NC>(a, X0, X1, Kp, hs, cth, sh);
      if (tid == 0) {
        const double fc = evfun(vm, s_par[5]);
        int done = 0;
        if (fabs(fc) < 1e-10 || (s_par[3] - s_par[1]) * hs < 1e-9) {
          s_par[1] = s_par[3] = cth;
          done = 1;
        } else if ((fc > 0.0) == (s_par[4] > 0.0)) {
          s_par[3] = cth; s_par[4] = fc;
          if (side == -1) s_par[2] *= 0.5;
          side = -1;
        } else {
          s_par[1] = cth; s_par[2] = fc;
          if (side == 1) s_par[4] *= 0.5;
          side = 1;
        }
        s_int[2] = done;
      }
      __syncthreads();
      if (s_int[2]) break;
    }
    const double th = s_par[3];
    // state at the crossing: statistics, onset snapshots, arrest drop and
    // the catalog sums over the ruptured cells (rsf_finalize_event,
    // src/rsf_output.c:1090-1135: |slip change|, change of |tau|, area
    // 4 l w, moment density area |slip change|)
    const int onset = !c->slipping;  // toggles to slipping
    double vm = 0.0, ss = 0.0, sm = 0.0, drop = 0.0, cnt = 0.0, y[D];
    double sds = 0.0, mds = 0.0, mdd = 0.0, sar = 0.0, smo = 0.0;
    for (int p = tid; p < n; p += blockDim.x) {
      dense_patch<NC>(a, X0, X1, Kp, hs, th, p, y);
      vm = fmax(vm, vel_mag<NC>(a.fp, a.coef[p], y));
      ss += y[NC + 2];
      sm += y[1] / y[NC + 1];
      if (onset) {
        a.tau0[p] = y[1];
        a.slip0[p] = y[NC + 2];
        a.rupt[p] = 0;
      } else if (a.rupt[p]) {
        const double dd = fabs(a.tau0[p]) - fabs(y[1]);
        const double ds = fabs(y[NC + 2] - a.slip0[p]);
        const double ar = a.area ? a.area[p] : 0.0;
        drop += dd;
        cnt += 1.0;
        sds += ds;
        mds = fmax(mds, ds);
        mdd = fmax(mdd, dd);
        sar += ar;
        smo += ar * ds;
      }
    }
    vm = block_reduce(vm, 1, sh);
    ss = block_reduce(ss, 0, sh);
    sm = block_reduce(sm, 0, sh);
    drop = block_reduce(drop, 0, sh);
    cnt = block_reduce(cnt, 0, sh);
    if (!onset) {
      sds = block_reduce(sds, 0, sh);
      mds = block_reduce(mds, 1, sh);
      mdd = block_reduce(mdd, 1, sh);
      sar = block_reduce(sar, 0, sh);
      smo = block_reduce(smo, 0, sh);
    }
    if (tid == 0) {
      const double tc = c->t + th * hs;
      const double peak = onset ? vm : fmax(c->peakv, vm);
      c->slipping = onset;
      if (onset) c->peakv = vm;
      if (c->nevent < c->max_event) {
        double* row = a.ev + 8 * c->nevent;
        row[0] = tc;
        row[1] = onset ? 1.0 : -1.0;
        row[2] = log10(vm > 1e-300 ? vm : 1e-300);
        row[3] = ss / n;
        row[4] = sm / n;
        row[5] = peak;
        row[6] = cnt > 0.0 ? drop / cnt : 0.0;
        row[7] = cnt;
        c->nevent++;
      } else {
        c->ev_drop++;
      }
      // catalog row of a closed event that ruptured at least one cell
      // (src/rsf_output.c:1150-1178): onset and arrest [s], ruptured
      // cells, their area, mean and max |slip change|, mean and max
      // stress drop [Pa], peak |v|, M0 = mu sum area |slip change|,
      // Mw = 2/3 (log10 M0 - 9.05) (-99 for M0 <= 0,
      // src/geometry/llgeo.c:757-760)
      if (!onset && c->ev_open && cnt > 0.0) {
        if (a.catr && c->ncat < c->max_event) {
          double* row = a.catr + kCatCols * c->ncat;
          const double m0 = o.shear_mu * smo;
          row[0] = c->t_onset;
          row[1] = tc;
          row[2] = cnt;
          row[3] = sar;
          row[4] = sds / cnt;
          row[5] = mds;
          row[6] = drop / cnt;
          row[7] = mdd;
          row[8] = peak;
          row[9] = m0;
          row[10] = (m0 > 0.0) ? (2.0 / 3.0) * (log10(m0) - 9.05) : -99.0;
          c->ncat++;
        } else {
          c->cat_drop++;
        }
      }
      if (onset) c->t_onset = tc;
      c->ev_open = onset;
      c->vprev = vm;
    }
    __syncthreads();
    th_lo = th;
    if (th_lo >= 1.0) {
      if (tid == 0) c->vprev = vnew;
      break;
    }
  }
  __syncthreads();
  // ---- advance
  if (tid == 0) {
    if (a.hlog && c->steps < a.hlog_cap) a.hlog[c->steps] = c->h;
    c->t += c->h;
    c->h = s_par[0];
    c->xcur = xo ^ 1;
    if (a.tb.imex) {
      c->imex_g1 = 1;  // stage 1 of the next step evaluates G at the new X
    } else {
      const int k0 = c->kslot[0];  // FSAL: K_{S-1} is K_0 of the next step
      c->kslot[0] = c->kslot[S - 1];
      c->kslot[S - 1] = k0;
    }
    c->steps++;
    c->attempt_failed = 0;
    c->last_vmax = vnew;
    if (c->slipping) {
      c->peakv = fmax(c->peakv, vnew);
      c->rupt_update = 1;
    }
    // monitor (rsf_TS_Monitor): change since the last logged state
    const double xn = sqrt(red[RED_XN2]), dx = sqrt(red[RED_DX2]);
    int bail = 0;
    if (xn > 1e-15 && dx / xn > o.rdx) bail = 1;
    if (o.adx > 0.0 && dx > o.adx) bail = 1;
    if (fabs(c->t - c->t_mon) > o.dt_monitor) bail = 1;
    if (bail) {
      if (c->nmon < c->max_mon) {
        double* row = a.mon + 10 * c->nmon;
        row[0] = (double)c->steps; row[1] = c->t; row[2] = c->t / kSecPerYear;
        row[3] = c->h; row[4] = log10(vnew); row[5] = red[RED_SLIP] / n;
        row[6] = red[RED_MU] / n; row[7] = red[RED_SMAX]; row[8] = red[RED_SMIN];
        const double mv = red[RED_VSUM] / n;
        row[9] = log10(mv > 1e-300 ? mv : 1e-300);
        c->nmon++;
      } else {
        c->mon_drop++;
      }
      c->t_mon = c->t;
      c->mon_copy = 1;
    }
    if (!(c->t < c->t_end)) c->halted = 1;
    if (c->steps >= c->max_steps) {
      c->halted = 1;
      c->status = 2;
    }
  }
}

// prep: first stage of an attempt (also Xmon copy and rupture marking)
template <typename VT, int NC>
__device__ __forceinline__ void prep_patch(const StageArgs& a, int p) {
  constexpr int D = Dim<NC>::v;
  const DevCtl* c = a.ctl;
  if (c->halted) return;
  const int n = a.n;
  if (p >= n) return;
  const int xc = c->xcur;
  const double* X = a.Xb[xc];
  const double* K0 = a.Stg[c->kslot[0]];
  const double h = c->h;
  const PatchCoef cf = load_coef(a.coef, n, p);
  double x[D], y[D];
#pragma unroll
  for (int f = 0; f < D; ++f) x[f] = X[f * n + p];
  if (c->mon_copy) {
#pragma unroll
    for (int f = 0; f < D; ++f) a.Xmon[f * n + p] = x[f];
  }
  if (c->rupt_update && vel_mag<NC>(a.fp, cf.a, x) >= a.so.rupture_vth) a.rupt[p] = 1;
  const double w = a.tb.A[1][0];
  // stage 1 is never the last stage of the supported pairs (s >= 4), so
  // its state is not stored: its state-only record carries it
#pragma unroll
  for (int f = 0; f < D; ++f) y[f] = x[f] + h * (w * K0[f * n + p]);
  double P[PreDim<NC>::v];
  rsf_patch_pre<NC>(a.fp, cf, y, P);
  store_pre<NC>(P, a.Pb[1], n, p);
  write_vel_pre<VT, NC>(P, static_cast<VT*>(a.vout), p, a.vpl);
  if (!rsf_patch_ok<NC>(a.fp, cf, y)) atomicOr(&a.ctl->stage_fail, 1);
}

// initial velocity field (before K_0 = F(X_0))
template <typename VT, int NC>
__device__ __forceinline__ void init_vel_patch(const StageArgs& a, int p) {
  constexpr int D = Dim<NC>::v;
  const int n = a.n;
  if (p >= n) return;
  const double* X = a.Xb[a.ctl->xcur];
  const PatchCoef cf = load_coef(a.coef, n, p);
  double x[D];
#pragma unroll
  for (int f = 0; f < D; ++f) x[f] = X[f * n + p];
  write_vel<VT, NC>(a.fp, cf.a, x, static_cast<VT*>(a.vout), p, a.vpl);
  if (!rsf_patch_ok<NC>(a.fp, cf, x)) atomicOr(&a.ctl->stage_fail, 1);
}

__device__ __forceinline__ double wave_reduce(double v, int op) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red_op(op, v, __shfl_xor(v, off, 64));
  return v;
}

// Step control after the final stage (s == S - 1) or after the initial
// evaluation (s == 0): one workgroup reduces the stage kernel's partials
// in a fixed order and runs the controller on an LDS copy of the control
// block, which it writes back at the end (one round trip to memory instead
// of a chain of dependent global accesses).
constexpr int kCtrlThreads = 256;

template <int NC>
__device__ __forceinline__ void ctrl_block(const StageArgs& a, int s, int nb) {
  __shared__ DevCtl sc;
  __shared__ double gred[kNRed];
  __shared__ double wred[kCtrlThreads / 64][kNRed];
  constexpr int NW = sizeof(DevCtl) / sizeof(long long);
  static_assert(sizeof(DevCtl) % sizeof(long long) == 0, "DevCtl copy granularity");
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // partials and the control block are fetched together (one memory round
  // trip); the partials are simply unused when the solve has halted
  double q[kNRed];
#pragma unroll
  for (int k = 0; k < kNRed; ++k) q[k] = red_identity(red_kind(k));
  for (int b = tid; b < nb; b += kCtrlThreads) {
#pragma unroll
    for (int k = 0; k < kNRed; ++k)
      q[k] = red_op(red_kind(k), q[k], a.partials[(size_t)k * nb + b]);
  }
  long long* cw = reinterpret_cast<long long*>(a.ctl);
  long long* sw = reinterpret_cast<long long*>(&sc);
  for (int i = tid; i < NW; i += kCtrlThreads) sw[i] = cw[i];
  // all nine reductions in one pass: shuffles within each wave, then the
  // wave results in a fixed order
#pragma unroll
  for (int k = 0; k < kNRed; ++k) q[k] = wave_reduce(q[k], red_kind(k));
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kNRed; ++k) wred[wave][k] = q[k];
  }
  __syncthreads();
  if (sc.halted) return;
  if (tid < kNRed) {
    double r = wred[0][tid];
    for (int w = 1; w < kCtrlThreads / 64; ++w) r = red_op(red_kind(tid), r, wred[w][tid]);
    gred[tid] = r;
  }
  __syncthreads();
  controller<NC>(a, s, gred, &sc);
  __syncthreads();
  for (int i = tid; i < NW; i += kCtrlThreads) cw[i] = sw[i];
}

// host: the argument block of a solve of handle h with operator K
void fill_stage_args(rsfg_ctx* h, const rsfg_solver_opts* o, const void* K, StageArgs& a);
// Butcher tableau of an integration method (RSFG_RK5DP, RSFG_RK3BS,
// RSFG_ARKIMEX_L2)
void fill_tableau(Tableau& tb, int method);
// stored operator of precision prec (nullptr if not assembled)
const void* operator_ptr(rsfg_ctx* h, int prec);
// elements of a product input vector for cols columns, zero-padded past the
// last column far enough for every product kernel's chunked reads
size_t vin_elems_for(int cols);

}  // namespace rsfg
