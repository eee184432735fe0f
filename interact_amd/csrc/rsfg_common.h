// Internal types shared by the librsfg translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rsfg.h"
#include "rsf_dev.h"

namespace rsfg {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define RSFG_HIP(call)                                                   \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) return ::rsfg::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define RSFG_CHECK(cond, code, msg) \
  do {                              \
    if (!(cond)) {                  \
      ::rsfg::set_error(msg);       \
      return (code);                \
    }                               \
  } while (0)

// Loads of operator entries: streamed once per product, so marked
// nontemporal; RSFG_TEMPORAL_K builds them as ordinary loads (a build for
// operators small enough to stay in the last-level cache between launches)
#ifndef RSFG_KLOAD
#ifdef RSFG_TEMPORAL_K
#define RSFG_KLOAD(...) (*(__VA_ARGS__))
#else
#define RSFG_KLOAD(...) __builtin_nontemporal_load(__VA_ARGS__)
#endif
#endif

// storage element of the bf16 hi+lo pair: hi in the upper 16 bits
typedef uint32_t bf16x2_t;
typedef uint16_t bf16_t;

// bf16 bits of an FP32 value, rounded to nearest even (finite values)
__device__ __forceinline__ uint32_t bf16_rn(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// capacity of the device event and monitor logs (rows)
constexpr int kMaxEv = 4096;
constexpr int kMaxMon = 1 << 18;
// maximum number of Runge-Kutta stages (rk5dp)
constexpr int kMaxStages = 7;
// reductions per workgroup in the final stage
constexpr int kNRed = 9;
constexpr int kCatCols = 11;   // doubles per catalog row

enum { RED_ERR = 0, RED_VMAX, RED_VSUM, RED_SLIP, RED_MU, RED_SMAX, RED_SMIN, RED_XN2, RED_DX2 };

// Controller state, resident in device memory; the single-workgroup
// control kernel updates it after every step attempt.
struct DevCtl {
  double t, h, t_end, hmax;
  double t_mon;
  double vprev;         // max|v| where the event function was last evaluated
  double peakv;         // running peak |v| since the last onset
  double last_vmax;
  long long steps, rejects, domain_rejects, rhs_evals, max_steps;
  int xcur;             // accepted-state buffer (0/1)
  int kslot[8];         // logical stage -> physical derivative slot
  int attempt_failed;   // previous attempt of this step was rejected
  int stage_fail;       // a stage domain check failed in this attempt
  int halted;
  int status;
  int slipping;
  int mon_copy;         // next prep copies X into Xmon
  int rupt_update;      // next prep marks ruptured cells
  int storm;
  int imex_g1;          // additive pair: the explicit part at X is stale
  int nevent, nmon;     // rows stored (at most max_event, max_mon)
  int max_event, max_mon;
  int ev_drop, mon_drop;  // rows past the capacity: counted, not stored
  double t_onset;       // time of the open event's onset
  int ev_open;          // an onset was logged and its arrest not yet
  int ncat, cat_drop;   // catalog rows stored (at most max_event) / past it
};

// Butcher tableau and integrator settings (kernel argument, by value)
struct Tableau {
  double A[kMaxStages][kMaxStages];   // explicit tableau
  double b[kMaxStages], bh[kMaxStages];
  double At[kMaxStages][kMaxStages];  // implicit tableau (additive pairs)
  double d1[kMaxStages], d2[kMaxStages];  // additive pairs: dense output
  int s, order, imex;
};

struct SolveOpts {
  double rtol, atol[6];
  double dt_monitor, rdx, adx;
  double vel_event, hyst, rupture_vth;
  double shear_mu;      // shear modulus [Pa] of the catalog's seismic moment
  int max_domain_reject;
};

}  // namespace rsfg

// the opaque handle of the C ABI
struct rsfg_ctx {
  int device = 0;
  int ncu = 0;                // compute units of the device (persistent kernels' grid)
  int n = 0;                  // patches
  bool borrowed = false;      // geometry and operator belong to another handle
  hipStream_t stream = nullptr;
  // geometry (device, SoA)
  double* d_geo = nullptr;    // kGeoFields x n, see kGeo*
  double* h_geo = nullptr;    // host copy
  // operator
  bool assembled = false;
  bool has_tri = false;        // any triangular patch (selects the assembly kernel)
  bool has_rect = false;       // any rectangular patch
  double tri_vtol = 5.0e-15;   // triangles' vertical-edge tolerance [rad] (the reference's)
  int slip_mode = 0, ncomp = 1, with_normal = 0, full_space = 0, store = 0;
  int G = 1;                  // rows per patch (ncomp shear + normal)
  int rows_all = 0;           // G * n
  int cols = 0;               // ncomp * n
  int ld = 0;                 // leading dimension (elements)
  double scale = 1.0;
  double* K64 = nullptr;
  float* K32 = nullptr;
  rsfg::bf16x2_t* Kbx2 = nullptr;
  rsfg::bf16_t* Kb16 = nullptr;
  size_t kslack = 0;          // extra elements allocated after the last row
  // bf16 + near-field lo format: hi in Kb16, per patch p the chunk list
  // nf_list[nf_off[p] .. nf_off[p+1]) and lo blocks from Knf_lo + nf_off[p] G 512
  rsfg::bf16_t* Knf_lo = nullptr;
  int* nf_off = nullptr;
  uint16_t* nf_list = nullptr;
  long nf_chunks = 0;
  double nf_tol = 0.0;
  // friction and loading
  bool have_friction = false, have_loading = false;
  rsfg::FricParams fp{};
  double* d_coef = nullptr;   // a, b, dc (3 n)
  double* d_sinc = nullptr;   // ncomp n shear loading + n normal loading
  // solver state
  int dim = 4;
  int xcur = 0;               // Xb buffer holding the accepted state
  double t = 0.0;
  bool have_state = false;
  double* d_state = nullptr;  // one block holding all solver vectors
  double *Xb[2] = {nullptr, nullptr}, *Yb[2] = {nullptr, nullptr};
  double* Stg[8] = {nullptr};
  double* Xmon = nullptr;
  double* Pb[2] = {nullptr, nullptr};   // state-only RHS records of stage states
  double* vin[2] = {nullptr, nullptr};  // D padded, used as double or float
  double* vxb[2] = {nullptr, nullptr};  // additive pair: v(X) per X buffer
  size_t vin_elems = 0;
  double* partials = nullptr;
  size_t partial_count = 0;
  double* d_ev = nullptr;
  double* d_mon = nullptr;
  double* d_cat = nullptr;    // tau0, slip0 per patch
  double* d_catr = nullptr;   // catalog rows (kCatCols each, ev_cap rows)
  double* d_area = nullptr;   // per-patch catalog area 4 l w (members borrow it)
  int* d_rupt = nullptr;
  rsfg::DevCtl* d_ctl = nullptr;
  rsfg::DevCtl* h_ctl = nullptr;  // pinned
  int nevent = 0, nmon = 0, ncat = 0;
  int ev_cap = 0, mon_cap = 0;    // rows allocated in d_ev (and d_catr), d_mon
  // the last solve can be continued (rsfg_solve_continue) with this method
  // and operator precision; cleared by every call that touches the buffers
  bool cont_ok = false;
  int cont_method = -1, cont_prec = 0;
  // captured graph of a batch of step attempts, reused while its argument
  // block (gkey) and batch size are unchanged
  hipGraph_t ggraph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;
  // accepted step sizes of the last solve (rsfg_set_step_log)
  double* d_hlog = nullptr;
  long long hlog_cap = 0;
  // scratch for API calls
  double* d_tmp = nullptr;
  size_t tmp_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace rsfg {
// geometry, structure of arrays: field f of patch p at d_geo[f * n + p]
enum {
  kGeoX = 0, kGeoY, kGeoZ,      // centroid
  kGeoCa, kGeoSa,               // cos, sin of alpha = 90 - strike
  kGeoSd, kGeoCd,               // sin, cos of dip after the 1e-8 snap of DC3D
  kGeoL, kGeoW,                 // half length, half width
  kGeoN0, kGeoN1, kGeoN2,       // normal
  kGeoS0, kGeoS1, kGeoS2,       // strike vector
  kGeoD0, kGeoD1, kGeoD2,       // dip vector
  kGeoR0, kGeoR1,               // rake fractions
  kGeoTri,                      // 1 for a triangle (vertices below), else 0
  kGeoV0,                       // triangle vertices: 9 fields x1 y1 z1 ... z3
  kGeoFields = kGeoV0 + 9
};

int ensure_tmp(rsfg_ctx* h, size_t bytes);
int alloc_state(rsfg_ctx* h);
// event and monitor logs of at least ev_rows and mon_rows rows
int ensure_logs(rsfg_ctx* h, int ev_rows, int mon_rows);
// drops the cached attempt graph (buffers moved or handle destroyed)
void drop_graph(rsfg_ctx* h);
}  // namespace rsfg
