// Discrete adjoint of the explicit Runge-Kutta cycle integration.
//
// The map differentiated is a replay: the rsfg_solve integration law
// (rsf_ODE_RHSFunction, src/rsf_engine.c:186-302, under the rk5dp / rk3bs
// tableaus of src/rsf_solve.c:440-546) advanced over a given list of step
// sizes, in practice the accepted steps of an adaptive solve (rsfg_set_step_log).
// With the steps fixed the replay is a smooth map of the friction
// parameters, and its adjoint is exact to rounding.
//
// Objective: J = 1/2 sum_n weight[n] sum_r (ln|v_r(X_n)| - target[n][r])^2
// over the replayed states X_0 .. X_N and the receiver patches r.
//
// Forward (one launch per stage): a wave per patch forms the elastic
// stressing rates K v(Y_i) of its rows, evaluates the patch right-hand side
// K_i = f(Y_i) and forms the next stage state and its slip rate.  The
// forward pass keeps a checkpoint every C steps (the state and the exact
// slip-rate vector it was evaluated with, so that recomputation is bitwise
// the same run).  Backward, segment by segment from the last: the segment's
// stage states and stressing rates are recomputed and recorded, then for
// each step, from the last stage to the first,
//   mu_i = h (b_i lambda' + sum_{l > i} a_li nu_l)
//   w_i  = (df/dtd)^T mu_i,  z_i = K^T w_i
//   nu_i = (df/ds)^T mu_i + (dv/ds)^T z_i,   g_p += (df/dp)^T mu_i + (dv/dp)^T z_i
//   lambda = lambda' + sum_i nu_i + dJ_n/dX_n
// The transposed local Jacobians come from forward-mode tangents through
// the pointwise law (adjoint_dev.h), one input direction at a time; K^T w
// is a column-strip product with row splits summed in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint_dev.h"
#include "stage_dev.h"

namespace rsfg {

constexpr int kRpWaves = 8;    // patches (waves) per workgroup of the replay stage kernel
constexpr int kAdjThreads = 256;

struct RpArgs {
  const void* K;
  long ld;
  int rows, cols, n;
  FricParams fp;
  const double* coef;  // a, b, dc
  const double* sinc;  // NC n shear loading + n normal loading
  const int* recv_of;  // receiver index of each patch, -1 if none
  int nrecv;
  double* trace;       // [(N + 1) x nrecv] ln|v| at the receivers
};

// one stage launch: evaluate Y (its slip rates are in vin), then form the
// next stage Yn = Xb + sum_j w[j] Kj[j] (w already multiplied by h) and its
// slip rates into vout.  Kj[j] == Kout takes the fresh derivative.
struct RpLaunch {
  const double* Y;
  double* Kout;
  double* recY;   // record of (Y, K v(Y)) or nullptr
  double* recT;
  long trow;      // trace row of Y, or -1
  const double* Xb;
  const double* Kj[kMaxStages];
  double w[kMaxStages];
  int nk;
  double* Yn;     // nullptr: nothing to form
  const void* vin;
  void* vout;
};

template <int NC>
__device__ __forceinline__ double vmag_of(const double* v) {
  return (NC == 1) ? fabs(v[0]) : sqrt(v[0] * v[0] + v[1] * v[1]);
}

template <typename TK, int NC>
__global__ __launch_bounds__(64 * kRpWaves) void k_rp_stage(RpArgs a, RpLaunch l) {
  using VT = typename KTraits<TK>::VT;
  constexpr int D = 2 * NC + 2;
  const int n = a.n;
  const int p = blockIdx.x * kRpWaves + (threadIdx.x >> 6);
  double td[NC];
  gemv_wave<TK, NC>(static_cast<const TK*>(a.K), a.ld, min(p, n - 1) * NC, a.rows, a.cols,
                    static_cast<const VT*>(l.vin), td);
  if ((threadIdx.x & 63) != 0 || p >= n) return;
  double y[D], k[D];
#pragma unroll
  for (int f = 0; f < D; ++f) y[f] = l.Y[(size_t)f * n + p];
  const double ca = a.coef[p], cb = a.coef[n + p], cd = a.coef[2 * n + p];
  adj::rhs<double, NC>(a.fp, ca, cb, cd, y, td, a.sinc + NC * p, a.sinc[NC * n + p], k);
#pragma unroll
  for (int f = 0; f < D; ++f) l.Kout[(size_t)f * n + p] = k[f];
  if (l.recY) {
#pragma unroll
    for (int f = 0; f < D; ++f) l.recY[(size_t)f * n + p] = y[f];
#pragma unroll
    for (int c = 0; c < NC; ++c) l.recT[(size_t)c * n + p] = td[c];
  }
  if (l.trow >= 0 && a.recv_of[p] >= 0) {
    double v[2];
    adj::vel<double, NC>(a.fp, ca, y, v);
    a.trace[l.trow * a.nrecv + a.recv_of[p]] = log(vmag_of<NC>(v));
  }
  if (l.Yn) {
    double yn[D];
#pragma unroll
    for (int f = 0; f < D; ++f) {
      double s = 0.0;
      for (int j = 0; j < l.nk; ++j)
        s += l.w[j] * ((l.Kj[j] == l.Kout) ? k[f] : l.Kj[j][(size_t)f * n + p]);
      yn[f] = l.Xb[(size_t)f * n + p] + s;
      l.Yn[(size_t)f * n + p] = yn[f];
    }
    double v[2];
    adj::vel<double, NC>(a.fp, ca, yn, v);
    VT* vo = static_cast<VT*>(l.vout);
#pragma unroll
    for (int c = 0; c < NC; ++c) vo[NC * p + c] = static_cast<VT>(v[c]);
  }
}

// slip rates of a state (the first product input of a replay)
template <typename VT, int NC>
__global__ __launch_bounds__(kAdjThreads) void k_rp_vel(RpArgs a, const double* Y, VT* v) {
  constexpr int D = 2 * NC + 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  double y[D], vv[2];
#pragma unroll
  for (int f = 0; f < D; ++f) y[f] = Y[(size_t)f * a.n + p];
  adj::vel<double, NC>(a.fp, a.coef[p], y, vv);
#pragma unroll
  for (int c = 0; c < NC; ++c) v[NC * p + c] = static_cast<VT>(vv[c]);
}

// J = 1/2 sum weight[n] (trace - target)^2, one workgroup, fixed order
__global__ __launch_bounds__(kAdjThreads) void k_objective(const double* trace,
                                                          const double* target,
                                                          const double* weight, long nstates,
                                                          int nrecv, double* J) {
  __shared__ double red[kAdjThreads];
  double s = 0.0;
  const long tot = nstates * nrecv;
  for (long i = threadIdx.x; i < tot; i += kAdjThreads) {
    const double r = trace[i] - target[i];
    s += 0.5 * weight[i / nrecv] * r * r;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kAdjThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *J = red[0];
}

// ---------------------------------------------------------------- backward
struct AdjArgs {
  int n, nsplit, cols, nrecv;
  FricParams fp;
  const double* coef;
  const double* sinc;
  double* M[kMaxStages];  // stage adjoint accumulators
  double* lam;            // D x n
  double* w;              // cols (row order of K)
  const double* zpart;    // nsplit x cols
  double* ga;
  double* gb;
  double* gdc;
  const int* recv_of;
  const double* target;
  const double* weight;
};

// (df/dtd)^T mu for one patch: the stressing-rate rows of the adjoint
template <int NC>
__device__ __forceinline__ void adj_w(const AdjArgs& A, int p, const double* y, const double* td,
                                      const double* mu, double* w) {
  using adj::Dual;
  constexpr int D = 2 * NC + 2;
  const int n = A.n;
  const Dual ca{A.coef[p], 0.0}, cb{A.coef[n + p], 0.0}, cd{A.coef[2 * n + p], 0.0};
  Dual s[D], out[D];
#pragma unroll
  for (int f = 0; f < D; ++f) s[f] = Dual{y[f], 0.0};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    Dual t[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) t[e] = Dual{td[e], (e == c) ? 1.0 : 0.0};
    adj::rhs<Dual, NC>(A.fp, ca, cb, cd, s, t, A.sinc + NC * p, A.sinc[NC * n + p], out);
    double acc = 0.0;
#pragma unroll
    for (int f = 0; f < D; ++f) acc += out[f].d * mu[f];
    w[c] = acc;
  }
}

// adds weight (ln|v| - target) d ln|v| / ds at the receiver patches to
// lam, and the same times d ln|v| / da to *ga
template <int NC>
__device__ __forceinline__ void adj_misfit(const AdjArgs& A, int p, long state, const double* y,
                                           double* lam, double* ga) {
  using adj::Dual;
  constexpr int D = 2 * NC + 2;
  const int r = A.recv_of[p];
  if (r < 0) return;
  const double wt = A.weight[state];
  if (wt == 0.0) return;
  double v0[2];
  adj::vel<double, NC>(A.fp, A.coef[p], y, v0);
  const double res = log(vmag_of<NC>(v0)) - A.target[state * A.nrecv + r];
  // directions psi, shear stresses, sigma, then a
#pragma unroll
  for (int q = 0; q < NC + 3; ++q) {
    const bool pa = (q == NC + 2);
    const Dual ca{A.coef[p], pa ? 1.0 : 0.0};
    Dual s[D], v[2];
#pragma unroll
    for (int f = 0; f < D; ++f) s[f] = Dual{y[f], (!pa && f == q) ? 1.0 : 0.0};
    adj::vel<Dual, NC>(A.fp, ca, s, v);
    const Dual vm = (NC == 1) ? adj::fabs(v[0]) : adj::sqrt(v[0] * v[0] + v[1] * v[1]);
    const double g = wt * res * (vm.d / vm.v);
    if (pa) *ga += g;
    else lam[q] += g;
  }
}

// lambda_N = dJ_N / dX_N
template <int NC>
__global__ __launch_bounds__(kAdjThreads) void k_adj_init(AdjArgs A, const double* X, long state) {
  constexpr int D = 2 * NC + 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const int n = A.n;
  double y[D], lam[D];
#pragma unroll
  for (int f = 0; f < D; ++f) {
    y[f] = X[(size_t)f * n + p];
    lam[f] = 0.0;
  }
  double ga = 0.0;
  adj_misfit<NC>(A, p, state, y, lam, &ga);
#pragma unroll
  for (int f = 0; f < D; ++f) A.lam[(size_t)f * n + p] = lam[f];
  A.ga[p] += ga;
}

// stage coefficients of one launch (h times a tableau row)
struct StageCoef {
  double c[kMaxStages];
};

// start of a step: M_j = h b_j lambda (hb.c[j], j < S - 1), then w of the
// last stage S - 2
template <int NC>
__global__ __launch_bounds__(kAdjThreads) void k_adj_begin(AdjArgs A, StageCoef hb, int S,
                                                         const double* recY,
                            const double* recT) {
  constexpr int D = 2 * NC + 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const int n = A.n;
  double lam[D], mu[D];
#pragma unroll
  for (int f = 0; f < D; ++f) lam[f] = A.lam[(size_t)f * n + p];
#pragma unroll
  for (int j = 0; j < kMaxStages - 1; ++j) {
    if (j < S - 1) {
#pragma unroll
      for (int f = 0; f < D; ++f) A.M[j][(size_t)f * n + p] = hb.c[j] * lam[f];
    }
    if (j == S - 2) {
#pragma unroll
      for (int f = 0; f < D; ++f) mu[f] = hb.c[j] * lam[f];
    }
  }
  double y[D], td[NC], w[NC];
#pragma unroll
  for (int f = 0; f < D; ++f) y[f] = recY[(size_t)f * n + p];
#pragma unroll
  for (int c = 0; c < NC; ++c) td[c] = recT[(size_t)c * n + p];
  adj_w<NC>(A, p, y, td, mu, w);
#pragma unroll
  for (int c = 0; c < NC; ++c) A.w[NC * p + c] = w[c];
}

// stage i of a step, given z_i = K^T w_i in row-split partials.  Eight
// consecutive lanes per patch, one input direction each: lanes 0 .. NC + 1
// the state components psi, the shear stresses and sigma (slip does not
// enter the law), lanes NC + 2 .. NC + 4 the parameters a, b, D_c.  The
// row splits of z are summed over the eight lanes and reduced in a fixed
// order, and the next stage's adjoint vector is exchanged by shuffles.
constexpr int kAdjLanes = 8;

template <int NC>
__global__ __launch_bounds__(kAdjThreads) void k_adj_stage(AdjArgs A, StageCoef ha, int i,
                                                         const double* recY,
                                                         const double* recT, const double* prevY,
                                                         const double* prevT, long state) {
  using adj::Dual;
  constexpr int D = 2 * NC + 2, NQ = NC + 2;
  static_assert(NQ + 3 <= kAdjLanes, "directions per patch");
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = gid / kAdjLanes, q = gid % kAdjLanes;
  const int base = (threadIdx.x & 63) - q;   // first lane of this patch's group
  const int n = A.n;
  const bool valid = p < n;
  const int pc = valid ? p : n - 1;
  // z: the row splits of K^T w, divided over the lanes, summed in order
  double z[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double s = 0.0;
    for (int k = q; k < A.nsplit; k += kAdjLanes) s += A.zpart[(size_t)k * A.cols + NC * pc + c];
#pragma unroll
    for (int off = kAdjLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kAdjLanes);
    z[c] = __shfl(s, base, 64);
  }
  const double* Mi = nullptr;
#pragma unroll
  for (int j = 0; j < kMaxStages - 1; ++j)
    if (j == i) Mi = A.M[j];
  double y[D], td[NC], mu[D];
#pragma unroll
  for (int f = 0; f < D; ++f) {
    y[f] = recY[(size_t)f * n + pc];
    mu[f] = Mi[(size_t)f * n + pc];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) td[c] = recT[(size_t)c * n + pc];
  const double a0 = A.coef[pc], b0 = A.coef[n + pc], d0 = A.coef[2 * n + pc];
  const double* sn = A.sinc + NC * pc;
  const double snn = A.sinc[NC * n + pc];
  // this lane's direction
  double g = 0.0;
  if (q < NQ + 3) {
    Dual s[D], t[NC], out[D], v[2];
#pragma unroll
    for (int f = 0; f < D; ++f) s[f] = Dual{y[f], (f == q) ? 1.0 : 0.0};
#pragma unroll
    for (int c = 0; c < NC; ++c) t[c] = Dual{td[c], 0.0};
    const Dual ca{a0, q == NQ ? 1.0 : 0.0}, cb{b0, q == NQ + 1 ? 1.0 : 0.0},
        cd{d0, q == NQ + 2 ? 1.0 : 0.0};
    adj::rhs<Dual, NC>(A.fp, ca, cb, cd, s, t, sn, snn, out);
#pragma unroll
    for (int f = 0; f < D; ++f) g += out[f].d * mu[f];
    if (q <= NQ) {   // the slip rates depend on the state and on a
      adj::vel<Dual, NC>(A.fp, ca, s, v);
#pragma unroll
      for (int c = 0; c < NC; ++c) g += v[c].d * z[c];
    }
  }
  // lanes q < NQ: nu_q; earlier stages M_j += h a_ij nu_i; lambda += nu_i
  double upd_prev = 0.0;
  if (q < NQ && valid) {
#pragma unroll
    for (int j = 0; j < kMaxStages - 1; ++j) {
      if (j < i && ha.c[j] != 0.0) A.M[j][(size_t)q * n + p] += ha.c[j] * g;
    }
  }
  if (q < NQ) {
    // the updated component q of M_{i-1} (what the lane just stored)
#pragma unroll
    for (int j = 0; j < kMaxStages - 1; ++j)
      if (i > 0 && j == i - 1) upd_prev = A.M[j][(size_t)q * n + pc];
  }
  double lam_q = 0.0;
  if (q < NQ) lam_q = A.lam[(size_t)q * n + pc] + g;
  double ga_add = (q == NQ) ? g : 0.0;
  if (i > 0) {
    // w of stage i - 1: mu' components from the lanes that own them (slip
    // components are unchanged by this stage and are loaded directly)
    double y1[D], t1[NC], mu1[D];
    const double* Mp = nullptr;
#pragma unroll
    for (int j = 0; j < kMaxStages - 1; ++j)
      if (j == i - 1) Mp = A.M[j];
#pragma unroll
    for (int f = 0; f < D; ++f) {
      y1[f] = prevY[(size_t)f * n + pc];
      const double own = __shfl(upd_prev, base + (f < NQ ? f : 0), 64);
      mu1[f] = (f < NQ) ? own : Mp[(size_t)f * n + pc];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) t1[c] = prevT[(size_t)c * n + pc];
    if (q < NC) {
      Dual s[D], t[NC], out[D];
#pragma unroll
      for (int f = 0; f < D; ++f) s[f] = Dual{y1[f], 0.0};
#pragma unroll
      for (int c = 0; c < NC; ++c) t[c] = Dual{t1[c], (c == q) ? 1.0 : 0.0};
      const Dual ca{a0, 0.0}, cb{b0, 0.0}, cd{d0, 0.0};
      adj::rhs<Dual, NC>(A.fp, ca, cb, cd, s, t, sn, snn, out);
      double w = 0.0;
#pragma unroll
      for (int f = 0; f < D; ++f) w += out[f].d * mu1[f];
      if (valid) A.w[NC * p + q] = w;
    }
  } else if (valid && A.recv_of[p] >= 0 && A.weight[state] != 0.0 && q <= NQ) {
    // the misfit of state X_n: d ln|v| / ds_q (lanes q < NQ) and / da (lane NQ)
    const double wt = A.weight[state];
    double v0v[2];
    adj::vel<double, NC>(A.fp, a0, y, v0v);
    const double res = log(vmag_of<NC>(v0v)) - A.target[state * A.nrecv + A.recv_of[p]];
    Dual s[D], v[2];
#pragma unroll
    for (int f = 0; f < D; ++f) s[f] = Dual{y[f], (f == q) ? 1.0 : 0.0};
    const Dual ca{a0, q == NQ ? 1.0 : 0.0};
    adj::vel<Dual, NC>(A.fp, ca, s, v);
    const Dual vm = (NC == 1) ? adj::fabs(v[0]) : adj::sqrt(v[0] * v[0] + v[1] * v[1]);
    const double dl = wt * res * (vm.d / vm.v);
    if (q < NQ) lam_q += dl;
    else ga_add += dl;
  }
  if (valid) {
    if (q < NQ) A.lam[(size_t)q * n + p] = lam_q;
    if (q == NQ) A.ga[p] += ga_add;
    if (q == NQ + 1) A.gb[p] += g;
    if (q == NQ + 2) A.gdc[p] += g;
  }
}

// zpart[split][col] = sum over the rows of this split of K[row][col] w[row];
// each thread owns 16 bytes of columns of every row of its split
template <typename TK>
__global__ __launch_bounds__(256) void k_adj_kt(const TK* __restrict__ K, long ld, int rows,
                                                int cols, const double* __restrict__ w,
                                                int rows_per_split, double* __restrict__ zpart) {
  constexpr int V = 16 / sizeof(TK);
  typedef TK vec_t __attribute__((ext_vector_type(V)));
  const int c0 = (blockIdx.x * 256 + threadIdx.x) * V;
  const int i0 = blockIdx.y * rows_per_split;
  const int i1 = min(rows, i0 + rows_per_split);
  if (c0 >= cols) return;
  double acc[V];
#pragma unroll
  for (int e = 0; e < V; ++e) acc[e] = 0.0;
  int i = i0;
  for (; i + 4 <= i1; i += 4) {
    vec_t kv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      kv[u] = RSFG_KLOAD(reinterpret_cast<const vec_t*>(K + (size_t)(i + u) * ld + c0));
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double wi = w[i + u];
#pragma unroll
      for (int e = 0; e < V; ++e) acc[e] = fma((double)kv[u][e], wi, acc[e]);
    }
  }
  for (; i < i1; ++i) {
    const vec_t kv = RSFG_KLOAD(reinterpret_cast<const vec_t*>(K + (size_t)i * ld + c0));
    const double wi = w[i];
#pragma unroll
    for (int e = 0; e < V; ++e) acc[e] = fma((double)kv[e], wi, acc[e]);
  }
#pragma unroll
  for (int e = 0; e < V; ++e)
    if (c0 + e < cols) zpart[(size_t)blockIdx.y * cols + c0 + e] = acc[e];
}

// ------------------------------------------------------------ host side

// device buffers of one replay / adjoint call
struct AdjWork {
  double* base = nullptr;
  double* Ys[kMaxStages] = {nullptr};
  double* Kd[kMaxStages] = {nullptr};
  double* vb[2] = {nullptr, nullptr};   // padded slip-rate vectors (VT)
  double* cpX = nullptr;                // checkpoints: state
  double* cpV = nullptr;                //   and slip-rate vector
  double* segY = nullptr;               // segment record
  double* segT = nullptr;
  double* lam = nullptr;
  double* M[kMaxStages] = {nullptr};
  double* w = nullptr;
  double* zpart = nullptr;
  double* g = nullptr;                  // ga, gb, gdc
  double* trace = nullptr;
  double* target = nullptr;
  double* weight = nullptr;
  double* J = nullptr;
  int* recv_of = nullptr;
  size_t vlen = 0;                      // elements of a padded vector (doubles)
  ~AdjWork() {
    hipFree(base);
    hipFree(recv_of);
  }
};

struct AdjSetup {
  rsfg_ctx* h;
  const rsfg_adjoint_problem* pb;
  Tableau tb;
  int S, D, NC, n, C, nsplit, rps;
  long long N;
  long ncp;
  const void* K;
  RpArgs ra;
};

static int adj_check(rsfg_ctx* h, const double* steps, long long nsteps,
                     const rsfg_adjoint_problem* pb, const char* who) {
  const std::string w(who);
  RSFG_CHECK(h && steps && pb && nsteps > 0, RSFG_EINVAL, w + ": bad arguments");
  RSFG_CHECK(h->assembled && h->have_friction && h->have_loading && h->have_state, RSFG_ESTATE,
             w + ": needs an assembled operator, friction, loading and a state");
  RSFG_CHECK(h->G == h->ncomp, RSFG_EINVAL,
             w + ": the operator holds normal rows (normal-stress coupling is not differentiated)");
  RSFG_CHECK(pb->method == RSFG_RK5DP || pb->method == RSFG_RK3BS, RSFG_EINVAL,
             w + ": method must be rk5dp or rk3bs");
  RSFG_CHECK(pb->prec == RSFG_PREC_FP64 || pb->prec == RSFG_PREC_FP32, RSFG_EINVAL,
             w + ": prec must be FP64 or FP32");
  RSFG_CHECK(operator_ptr(h, pb->prec) != nullptr, RSFG_ESTATE,
             w + ": operator precision not assembled");
  RSFG_CHECK(pb->nrecv >= 0 && (pb->nrecv == 0 || pb->recv), RSFG_EINVAL,
             w + ": receivers missing");
  for (int r = 0; r < pb->nrecv; ++r)
    RSFG_CHECK(pb->recv[r] >= 0 && pb->recv[r] < h->n, RSFG_EINVAL,
               w + ": receiver index out of range");
  for (long long k = 0; k < nsteps; ++k)
    RSFG_CHECK(steps[k] > 0.0 && std::isfinite(steps[k]), RSFG_EINVAL,
               w + ": step sizes must be positive and finite");
  return RSFG_OK;
}

static int adj_alloc(AdjSetup& s, AdjWork& W, bool gradient) {
  rsfg_ctx* h = s.h;
  const int n = h->n, D = s.D, S = s.S;
  const size_t st = (size_t)D * n;
  W.vlen = vin_elems_for(h->cols);
  const long long nstates = s.N + 1;
  const int nrecv = s.pb->nrecv;
  size_t tot = 0;
  tot += 2 * (size_t)S * st;                      // Ys, Kd
  tot += 2 * W.vlen;                              // vb
  tot += (size_t)s.ncp * (st + W.vlen);           // checkpoints
  tot += (size_t)nstates * std::max(nrecv, 1);    // trace
  tot += (size_t)nstates * std::max(nrecv, 1);    // target
  tot += (size_t)nstates;                         // weight
  tot += 1;                                       // J
  if (gradient) {
    tot += (size_t)s.C * (S - 1) * (st + (size_t)s.NC * n);  // segment record
    tot += st * S;                                // lam, M
    tot += h->cols;                               // w
    tot += (size_t)s.nsplit * h->cols;            // zpart
    tot += 3 * (size_t)n;                         // grads
  }
  RSFG_HIP(hipMalloc(&W.base, tot * sizeof(double)));
  RSFG_HIP(hipMalloc(&W.recv_of, sizeof(int) * n));
  double* q = W.base;
  auto take = [&](size_t k) {
    double* r = q;
    q += k;
    return r;
  };
  for (int i = 0; i < S; ++i) W.Ys[i] = take(st);
  for (int i = 0; i < S; ++i) W.Kd[i] = take(st);
  W.vb[0] = take(W.vlen);
  W.vb[1] = take(W.vlen);
  W.cpX = take((size_t)s.ncp * st);
  W.cpV = take((size_t)s.ncp * W.vlen);
  W.trace = take((size_t)nstates * std::max(nrecv, 1));
  W.target = take((size_t)nstates * std::max(nrecv, 1));
  W.weight = take((size_t)nstates);
  W.J = take(1);
  if (gradient) {
    W.segY = take((size_t)s.C * (S - 1) * st);
    W.segT = take((size_t)s.C * (S - 1) * s.NC * n);
    W.lam = take(st);
    for (int j = 0; j < S - 1; ++j) W.M[j] = take(st);
    W.w = take(h->cols);
    W.zpart = take((size_t)s.nsplit * h->cols);
    W.g = take(3 * (size_t)n);
  }
  // the product reads the slip-rate vectors a chunk past the last column
  RSFG_HIP(hipMemsetAsync(W.vb[0], 0, 2 * W.vlen * sizeof(double), h->stream));
  RSFG_HIP(hipMemsetAsync(W.cpV, 0, (size_t)s.ncp * W.vlen * sizeof(double), h->stream));
  std::vector<int> ro(n, -1);
  for (int r = 0; r < nrecv; ++r) ro[s.pb->recv[r]] = r;
  RSFG_HIP(hipMemcpyAsync(W.recv_of, ro.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
  if (nrecv > 0 && s.pb->target)
    RSFG_HIP(hipMemcpyAsync(W.target, s.pb->target, sizeof(double) * nstates * nrecv,
                            hipMemcpyHostToDevice, h->stream));
  if (s.pb->weight)
    RSFG_HIP(hipMemcpyAsync(W.weight, s.pb->weight, sizeof(double) * nstates,
                            hipMemcpyHostToDevice, h->stream));
  else
    RSFG_HIP(hipMemsetAsync(W.weight, 0, sizeof(double) * nstates, h->stream));
  RSFG_HIP(hipMemsetAsync(W.trace, 0, sizeof(double) * nstates * std::max(nrecv, 1), h->stream));
  return RSFG_OK;
}

template <typename TK, int NC>
static void launch_rp(const AdjSetup& s, const RpLaunch& l) {
  const int nb = (s.n + kRpWaves - 1) / kRpWaves;
  hipLaunchKernelGGL((k_rp_stage<TK, NC>), dim3(nb), dim3(64 * kRpWaves), 0, s.h->stream, s.ra, l);
}

static void rp(const AdjSetup& s, const RpLaunch& l) {
  const bool f64 = s.pb->prec == RSFG_PREC_FP64;
  if (s.NC == 2) f64 ? launch_rp<double, 2>(s, l) : launch_rp<float, 2>(s, l);
  else f64 ? launch_rp<double, 1>(s, l) : launch_rp<float, 1>(s, l);
}

// Replay steps [s0, s1) from the state in Ys[0] whose slip rates are in
// vb[*vc].  Steps in [r0, r1) record their stage states and stressing rates
// into the segment record (step r0 first); trace: ln|v| rows; cp:
// checkpoints at multiples of C.
static void replay_steps(AdjSetup& s, AdjWork& W, const double* steps, long long s0,
                         long long s1, bool first_eval, long long r0, long long r1, bool trace,
                         bool cp, int* vc) {
  rsfg_ctx* h = s.h;
  const int S = s.S, n = s.n;
  const size_t st = (size_t)s.D * n;
  const size_t vbytes = W.vlen * sizeof(double);
  auto rec_of = [&](long long step, int stage, double** ry, double** rt) {
    if (step < r0 || step >= r1) return;
    const size_t k = (size_t)(step - r0) * (S - 1) + stage;
    *ry = W.segY + k * st;
    *rt = W.segT + k * (size_t)s.NC * n;
  };
  // evaluation of X_{s0} and the first stage of step s0
  if (first_eval) {
    RpLaunch l;
    memset(&l, 0, sizeof l);
    l.Y = W.Ys[0];
    l.Kout = W.Kd[0];
    rec_of(s0, 0, &l.recY, &l.recT);
    l.trow = trace ? s0 : -1;
    l.Xb = W.Ys[0];
    l.Kj[0] = W.Kd[0];
    l.w[0] = steps[s0] * s.tb.A[1][0];
    l.nk = 1;
    l.Yn = W.Ys[1];
    l.vin = W.vb[*vc];
    l.vout = W.vb[*vc ^ 1];
    rp(s, l);
    *vc ^= 1;
  }
  for (long long m = s0; m < s1; ++m) {
    const double hm = steps[m];
    for (int i = 1; i < S; ++i) {
      RpLaunch l;
      memset(&l, 0, sizeof l);
      l.Y = W.Ys[i];
      l.Kout = W.Kd[i];
      l.vin = W.vb[*vc];
      l.vout = W.vb[*vc ^ 1];
      if (i < S - 1) {
        rec_of(m, i, &l.recY, &l.recT);
        l.trow = -1;
        l.Xb = W.Ys[0];
        for (int j = 0; j <= i; ++j) {
          l.Kj[j] = W.Kd[j];
          l.w[j] = hm * s.tb.A[i + 1][j];
        }
        l.nk = i + 1;
        l.Yn = W.Ys[i + 1];
      } else {
        // Y_{S-1} = X_{m+1}: its derivative is K_0 of step m + 1
        if (m + 1 < s1) rec_of(m + 1, 0, &l.recY, &l.recT);
        l.trow = trace ? m + 1 : -1;
        if (m + 1 < s.N) {
          l.Xb = W.Ys[S - 1];
          l.Kj[0] = W.Kd[S - 1];
          l.w[0] = steps[m + 1] * s.tb.A[1][0];
          l.nk = 1;
          l.Yn = W.Ys[1];
        }
      }
      rp(s, l);
      *vc ^= 1;
    }
    std::swap(W.Ys[0], W.Ys[S - 1]);
    std::swap(W.Kd[0], W.Kd[S - 1]);
    // checkpoint of X_{m+1} with the slip rates it is evaluated with
    if (cp && (m + 1) % s.C == 0 && m + 1 < s.N) {
      const long k = (long)((m + 1) / s.C);
      hipMemcpyAsync(W.cpX + k * st, W.Ys[0], st * sizeof(double), hipMemcpyDeviceToDevice,
                     h->stream);
      // (the last launch turned the parity: vb[*vc] holds v of the next
      // step's first stage, vb[*vc ^ 1] the slip rates of X_{m+1})
      hipMemcpyAsync(W.cpV + k * W.vlen, W.vb[*vc ^ 1], vbytes, hipMemcpyDeviceToDevice,
                     h->stream);
    }
  }
}

static int adj_setup(rsfg_ctx* h, const double* steps, long long nsteps,
                     const rsfg_adjoint_problem* pb, AdjSetup& s) {
  (void)steps;
  s.h = h;
  s.pb = pb;
  fill_tableau(s.tb, pb->method);
  s.S = s.tb.s;
  s.NC = h->ncomp;
  s.D = h->dim;
  s.n = h->n;
  s.N = nsteps;
  s.C = pb->checkpoint_every > 0 ? pb->checkpoint_every : 256;
  s.ncp = (long)((nsteps + s.C - 1) / s.C);
  s.K = operator_ptr(h, pb->prec);
  const int V = (pb->prec == RSFG_PREC_FP64) ? 2 : 4;
  const int strips = (h->cols + 256 * V - 1) / (256 * V);
  s.nsplit = std::max(1, std::min(std::max(1, 2048 / strips), std::max(1, h->rows_all / 8)));
  s.rps = (h->rows_all + s.nsplit - 1) / s.nsplit;
  s.nsplit = (h->rows_all + s.rps - 1) / s.rps;
  memset(&s.ra, 0, sizeof s.ra);
  s.ra.K = s.K;
  s.ra.ld = h->ld;
  s.ra.rows = h->rows_all;
  s.ra.cols = h->cols;
  s.ra.n = h->n;
  s.ra.fp = h->fp;
  s.ra.coef = h->d_coef;
  s.ra.sinc = h->d_sinc;
  s.ra.nrecv = pb->nrecv;
  return RSFG_OK;
}

// first pass: X_0 from the handle, trace, checkpoints, J; with a gradient
// to follow (record_last) the stages of the last segment are recorded on
// the way, so that the backward sweep starts without recomputing them
static int forward_pass(AdjSetup& s, AdjWork& W, const double* steps, int* vc,
                        bool record_last) {
  rsfg_ctx* h = s.h;
  const size_t st = (size_t)s.D * s.n;
  s.ra.recv_of = W.recv_of;
  s.ra.trace = W.trace;
  RSFG_HIP(hipMemcpyAsync(W.Ys[0], h->Xb[h->xcur], st * sizeof(double), hipMemcpyDeviceToDevice,
                          h->stream));
  const dim3 gp((s.n + kAdjThreads - 1) / kAdjThreads);
  *vc = 0;
  if (s.pb->prec == RSFG_PREC_FP64) {
    if (s.NC == 2) hipLaunchKernelGGL((k_rp_vel<double, 2>), gp, dim3(kAdjThreads), 0, h->stream, s.ra, W.Ys[0], W.vb[0]);
    else hipLaunchKernelGGL((k_rp_vel<double, 1>), gp, dim3(kAdjThreads), 0, h->stream, s.ra, W.Ys[0], W.vb[0]);
  } else {
    float* v = reinterpret_cast<float*>(W.vb[0]);
    if (s.NC == 2) hipLaunchKernelGGL((k_rp_vel<float, 2>), gp, dim3(kAdjThreads), 0, h->stream, s.ra, W.Ys[0], v);
    else hipLaunchKernelGGL((k_rp_vel<float, 1>), gp, dim3(kAdjThreads), 0, h->stream, s.ra, W.Ys[0], v);
  }
  RSFG_HIP(hipMemcpyAsync(W.cpX, W.Ys[0], st * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  RSFG_HIP(hipMemcpyAsync(W.cpV, W.vb[0], W.vlen * sizeof(double), hipMemcpyDeviceToDevice,
                          h->stream));
  const long long last0 = (long long)(s.ncp - 1) * s.C;
  replay_steps(s, W, steps, 0, s.N, true, record_last ? last0 : 0, record_last ? s.N : 0, true,
               true, vc);
  if (s.pb->nrecv > 0)
    hipLaunchKernelGGL(k_objective, dim3(1), dim3(kAdjThreads), 0, h->stream, W.trace, W.target,
                       W.weight, (long)(s.N + 1), s.pb->nrecv, W.J);
  else
    RSFG_HIP(hipMemsetAsync(W.J, 0, sizeof(double), h->stream));
  RSFG_HIP(hipGetLastError());
  return RSFG_OK;
}

template <typename TK>
static void launch_kt(const AdjSetup& s, const AdjWork& W) {
  constexpr int V = 16 / sizeof(TK);
  const dim3 grid((s.h->cols + 256 * V - 1) / (256 * V), s.nsplit);
  hipLaunchKernelGGL(k_adj_kt<TK>, grid, dim3(256), 0, s.h->stream, static_cast<const TK*>(s.K),
                     (long)s.h->ld, s.h->rows_all, s.h->cols, W.w, s.rps, W.zpart);
}

template <int NC>
static int backward_pass(AdjSetup& s, AdjWork& W, const double* steps) {
  rsfg_ctx* h = s.h;
  const int n = s.n, S = s.S;
  const size_t st = (size_t)s.D * n;
  AdjArgs A;
  memset(&A, 0, sizeof A);
  A.n = n;
  A.nsplit = s.nsplit;
  A.cols = h->cols;
  A.nrecv = s.pb->nrecv;
  A.fp = h->fp;
  A.coef = h->d_coef;
  A.sinc = h->d_sinc;
  for (int j = 0; j < S - 1; ++j) A.M[j] = W.M[j];
  A.lam = W.lam;
  A.w = W.w;
  A.zpart = W.zpart;
  A.ga = W.g;
  A.gb = W.g + n;
  A.gdc = W.g + 2 * n;
  A.recv_of = W.recv_of;
  A.target = W.target;
  A.weight = W.weight;
  const dim3 gp((n + kAdjThreads - 1) / kAdjThreads);
  const dim3 gp8(((size_t)n * kAdjLanes + kAdjThreads - 1) / kAdjThreads);
  RSFG_HIP(hipMemsetAsync(W.g, 0, 3 * (size_t)n * sizeof(double), h->stream));
  // lambda_N from the final state (Ys[0] after the forward pass)
  hipLaunchKernelGGL(k_adj_init<NC>, gp, dim3(kAdjThreads), 0, h->stream, A, W.Ys[0], (long)s.N);
  const bool f64 = s.pb->prec == RSFG_PREC_FP64;
  for (long k = s.ncp - 1; k >= 0; --k) {
    const long long s0 = (long long)k * s.C, s1 = std::min<long long>(s.N, s0 + s.C);
    // recompute the segment from its checkpoint, recording every stage (the
    // last segment was recorded by the forward pass)
    if (k < s.ncp - 1) {
      RSFG_HIP(hipMemcpyAsync(W.Ys[0], W.cpX + k * st, st * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
      RSFG_HIP(hipMemcpyAsync(W.vb[0], W.cpV + k * W.vlen, W.vlen * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
      int vc = 0;
      replay_steps(s, W, steps, s0, s1, true, s0, s1, false, false, &vc);
    }
    for (long long m = s1 - 1; m >= s0; --m) {
      const double hm = steps[m];
      const size_t base = (size_t)(m - s0) * (S - 1);
      auto recY = [&](int i) { return W.segY + (base + i) * st; };
      auto recT = [&](int i) { return W.segT + (base + i) * (size_t)NC * n; };
      StageCoef hb;
      memset(&hb, 0, sizeof hb);
      for (int j = 0; j < S - 1; ++j) hb.c[j] = hm * s.tb.A[S - 1][j];
      hipLaunchKernelGGL(k_adj_begin<NC>, gp, dim3(kAdjThreads), 0, h->stream, A, hb, S, recY(S - 2),
                         recT(S - 2));
      for (int i = S - 2; i >= 0; --i) {
        f64 ? launch_kt<double>(s, W) : launch_kt<float>(s, W);
        const double* py = i > 0 ? recY(i - 1) : nullptr;
        const double* pt = i > 0 ? recT(i - 1) : nullptr;
        StageCoef ha;
        memset(&ha, 0, sizeof ha);
        for (int j = 0; j < i; ++j) ha.c[j] = hm * s.tb.A[i][j];
        hipLaunchKernelGGL(k_adj_stage<NC>, gp8, dim3(kAdjThreads), 0, h->stream, A, ha, i,
                           recY(i), recT(i), py, pt, (long)m);
      }
    }
    RSFG_HIP(hipGetLastError());
  }
  return RSFG_OK;
}

}  // namespace rsfg

using namespace rsfg;

extern "C" {

int rsfg_replay(rsfg_handle* h, const double* steps, long long nsteps,
                const rsfg_adjoint_problem* pb, double* J, double* trace, double* X_end) {
  if (h) h->cont_ok = false;
  int rc = adj_check(h, steps, nsteps, pb, "rsfg_replay");
  if (rc) return rc;
  RSFG_HIP(hipSetDevice(h->device));
  RSFG_CHECK(!J || (pb->nrecv > 0 && pb->target && pb->weight), RSFG_EINVAL,
             "rsfg_replay: the objective needs receivers, a target and weights");
  AdjSetup s;
  adj_setup(h, steps, nsteps, pb, s);
  AdjWork W;
  rc = adj_alloc(s, W, false);
  if (rc) return rc;
  int vc = 0;
  rc = forward_pass(s, W, steps, &vc, false);
  if (rc) return rc;
  if (J) RSFG_HIP(hipMemcpyAsync(J, W.J, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (trace && pb->nrecv > 0)
    RSFG_HIP(hipMemcpyAsync(trace, W.trace, sizeof(double) * (nsteps + 1) * pb->nrecv,
                            hipMemcpyDeviceToHost, h->stream));
  std::vector<double> soa;
  if (X_end) {
    soa.resize((size_t)s.D * s.n);
    RSFG_HIP(hipMemcpyAsync(soa.data(), W.Ys[0], sizeof(double) * soa.size(),
                            hipMemcpyDeviceToHost, h->stream));
  }
  RSFG_HIP(hipStreamSynchronize(h->stream));
  if (X_end)
    for (int p = 0; p < s.n; ++p)
      for (int f = 0; f < s.D; ++f) X_end[(size_t)p * s.D + f] = soa[(size_t)f * s.n + p];
  return RSFG_OK;
}

int rsfg_adjoint_gradient(rsfg_handle* h, const double* steps, long long nsteps,
                          const rsfg_adjoint_problem* pb, double* J, double* grad_a,
                          double* grad_b, double* grad_dc) {
  if (h) h->cont_ok = false;
  int rc = adj_check(h, steps, nsteps, pb, "rsfg_adjoint_gradient");
  if (rc) return rc;
  RSFG_CHECK(pb->nrecv > 0 && pb->target && pb->weight, RSFG_EINVAL,
             "rsfg_adjoint_gradient: needs receivers, a target and weights");
  RSFG_HIP(hipSetDevice(h->device));
  AdjSetup s;
  adj_setup(h, steps, nsteps, pb, s);
  AdjWork W;
  rc = adj_alloc(s, W, true);
  if (rc) return rc;
  int vc = 0;
  rc = forward_pass(s, W, steps, &vc, true);
  if (rc) return rc;
  rc = (s.NC == 2) ? backward_pass<2>(s, W, steps) : backward_pass<1>(s, W, steps);
  if (rc) return rc;
  const int n = s.n;
  if (J) RSFG_HIP(hipMemcpyAsync(J, W.J, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (grad_a)
    RSFG_HIP(hipMemcpyAsync(grad_a, W.g, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (grad_b)
    RSFG_HIP(hipMemcpyAsync(grad_b, W.g + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (grad_dc)
    RSFG_HIP(hipMemcpyAsync(grad_dc, W.g + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost,
                            h->stream));
  RSFG_HIP(hipStreamSynchronize(h->stream));
  return RSFG_OK;
}

int rsfg_adjoint_bench(rsfg_handle* h, int prec, int nrep, double* ms) {
  if (h) h->cont_ok = false;
  RSFG_CHECK(h && ms && nrep > 0, RSFG_EINVAL, "rsfg_adjoint_bench: bad arguments");
  RSFG_CHECK(h->assembled && h->G == h->ncomp, RSFG_ESTATE,
             "rsfg_adjoint_bench: needs an assembled operator without normal rows");
  RSFG_CHECK(prec == RSFG_PREC_FP64 || prec == RSFG_PREC_FP32, RSFG_EINVAL,
             "rsfg_adjoint_bench: prec must be FP64 or FP32");
  RSFG_CHECK(operator_ptr(h, prec) != nullptr, RSFG_ESTATE,
             "rsfg_adjoint_bench: operator precision not assembled");
  RSFG_HIP(hipSetDevice(h->device));
  rsfg_adjoint_problem pb;
  memset(&pb, 0, sizeof pb);
  pb.method = RSFG_RK5DP;
  pb.prec = prec;
  const double one = 1.0;
  AdjSetup s;
  adj_setup(h, &one, 1, &pb, s);
  AdjWork W;
  RSFG_HIP(hipMalloc(&W.base, sizeof(double) * (h->cols + (size_t)s.nsplit * h->cols)));
  W.w = W.base;
  W.zpart = W.base + h->cols;
  RSFG_HIP(hipMemsetAsync(W.w, 0, sizeof(double) * h->cols, h->stream));
  const bool f64 = prec == RSFG_PREC_FP64;
  f64 ? launch_kt<double>(s, W) : launch_kt<float>(s, W);
  RSFG_HIP(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < nrep; ++r) f64 ? launch_kt<double>(s, W) : launch_kt<float>(s, W);
  RSFG_HIP(hipEventRecord(h->ev1, h->stream));
  RSFG_HIP(hipEventSynchronize(h->ev1));
  RSFG_HIP(hipGetLastError());
  float t = 0.f;
  RSFG_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t / nrep;
  return RSFG_OK;
}

}  // extern "C"
