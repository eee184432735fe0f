/*
  librsfg -- MI355X (gfx950) implementation of the rsf_solve dense
  quasi-dynamic cycle hot path, behind a plain C ABI.

  The reference (thwbecker/interact) runs this path on the CPU through PETSc:
  operator assembly in calc_petsc_Isn_matrices / GenKEntries_petsc, the
  stress product MatMult(Is, v), the RHS callback rsf_ODE_RHSFunction with
  rsf_domain_check, and the TSRK/TSAdapt/TSEvent loop with the rsf monitor.
  Each entry point below names the reference interface it replaces.

  Conventions
   - every function returns 0 (RSFG_OK) or a negative error code; nothing
     calls exit(); rsfg_last_error() gives a message;
   - host pointers unless a name ends in _dev (device pointers on the
     handle's GPU);
   - a handle owns one device, one HIP stream and all device memory it
     allocates; handles are independent (several may be driven from
     separate host threads); calls on one handle must not overlap;
   - state vectors use the reference's patch-major layout
     (src/rsf_solve.c:364-372): scalar slip mode (psi, tau, sigma, slip)
     per patch; two-component mode (psi, tau_s, tau_d, sigma, slip_s,
     slip_d) per patch.
*/
#ifndef RSFG_H
#define RSFG_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RSFG_OK 0
#define RSFG_EINVAL -1   /* bad argument */
#define RSFG_ENOMEM -2   /* device or host allocation failed */
#define RSFG_EHIP -3     /* HIP runtime error */
#define RSFG_ESTATE -4   /* call out of order (e.g. solve before assemble) */
#define RSFG_ESTORM -5   /* domain-check rejection storm (src/rsf_engine.c:336-358) */
#define RSFG_EDOMAIN -6  /* non-finite or out-of-domain accepted state */

/* slip modes, as the reference's MODE_TYPE STRIKE/DIP/RAKE
   (-rsf_slip_mode, src/rsf_init.c:459-468) plus the two-component mode */
#define RSFG_SLIP_STRIKE 0
#define RSFG_SLIP_DIP 1
#define RSFG_SLIP_RAKE 3
#define RSFG_SLIP_VECTOR 8   /* strike and dip slip both free: 2 DOF/patch */

/* storage precisions of the operator K (bit mask for rsfg_assemble) */
#define RSFG_PREC_FP64 1     /* master copy, always kept */
#define RSFG_PREC_FP32 2     /* FP32 entries */
#define RSFG_PREC_BF16X2 4   /* bf16 hi + bf16 lo pair (hi+lo = FP32 value
                                rounded to 16 significant bits) */
#define RSFG_PREC_BF16 8     /* bf16 hi only */
#define RSFG_PREC_BF16NF 16  /* bf16 hi for every entry plus the bf16 lo of the
                                BF16X2 pair in each patch's near-field column
                                chunks (512 columns), chosen per patch so that
                                the chunks left without lo hold at most a
                                fraction nf_tol^2 of the patch rows' energy
                                (sum of squared entries) */

/* time integrators (PETSc -ts_rk_type 5dp / 3bs; -imex selects PETSc
   TSARKIMEX with tableau l2, src/rsf_imex.c:257-305) */
#define RSFG_RK5DP 0
#define RSFG_RK3BS 1
#define RSFG_ARKIMEX_L2 2  /* additive IMEX: stressing rates explicit, local
                              state evolution implicit by per-cell Newton
                              (src/rsf_imex.c); scalar slip layout only */

/* one fault patch as in the reference's patch file
   "x y z strike dip L W group" (src/read_geometry.c:115-117, strike and
   dip kept as float like struct flt, src/includes/structures.h:69).
   L < 0 and W < 0 mark a triangle whose three vertices follow on the
   line (src/read_geometry.c:148-215): vert holds them, x y z strike dip
   are then ignored (the receiver point is the vertex mean and the plane
   vectors are the triangle's TDCS vectors). */
typedef struct {
  double x, y, z;       /* centroid [m], z <= 0 in a half space */
  float strike, dip;    /* degrees */
  double l, w;          /* half length, half width [m]; both < 0: triangle */
  double rake[2];       /* (strike, dip) fractions for RSFG_SLIP_RAKE */
  int group;
  double vert[9];       /* triangle vertices x1 y1 z1 x2 ... z3 [m], z <= 0 */
} rsfg_patch;

typedef struct rsfg_ctx rsfg_handle;

const char *rsfg_last_error(void);
/* library version string */
const char *rsfg_version(void);

/* number of visible devices (0 without a GPU) */
int rsfg_device_count(int *count);

/* Binds patches to a device and uploads the geometry (replaces the patch
   part of read_geometry, src/read_geometry.c:23-700, whose per-patch
   frames are recomputed here in FP64 from the float angles). */
int rsfg_create(int device, const rsfg_patch *patches, int n,
                rsfg_handle **h);
int rsfg_destroy(rsfg_handle *h);

/* Reads a reference patch file (read_geometry, src/read_geometry.c:115-117;
   rectangles, and triangles with L, W < 0 followed by their nine vertex
   coordinates, :148-215).  On entry *n is the capacity of patches; on
   return the number of patches the file holds.  patches may be NULL (with
   any capacity) to count them.  A file holding more patches than the
   capacity returns RSFG_EINVAL after storing the first *n-in patches, with
   *n set to the full count, so the caller can retry with enough room (the
   reference reads every patch). */
int rsfg_read_patch_file(const char *fname, rsfg_patch *patches, int *n);

/* A handle that shares geom's patches and assembled operators (read only)
   and owns its own friction, loading, state and solver buffers: members of
   an ensemble over one fault geometry (no counterpart in the reference,
   which runs one parameter set per process).  Destroy members before their
   geometry; a member cannot re-assemble. */
int rsfg_create_member(rsfg_handle *geom, rsfg_handle **member);

/* ---- operator assembly ----------------------------------------------
   Replaces the dense branch of calc_petsc_Isn_matrices
   (src/petsc_matrix_operations.c:379-503, MatScale at :735) and
   GenKEntries_petsc (src/interact_matrix_assembly.c:102-151): one GPU
   thread per (receiver, source) pair evaluates the Okada (1992) stress at
   the receiver centroid for unit strike and unit dip slip on the source,
   resolves tractions on the receiver plane, scales, and stores.
   Rows are receivers, columns sources.  Scalar slip modes give the N x N
   "Is" operator; RSFG_SLIP_VECTOR gives 2N x 2N with dof 2p+c, c = 0
   strike, 1 dip.  with_normal adds the "In" rows (normal traction,
   compression positive, i.e. scaled by -scale, src/rsf_solve.c:245-247).
   scale is G_SI / 1e4 (src/rsf_solve.c:235-237).  Singular pairs store 0. */
typedef struct {
  int slip_mode;        /* RSFG_SLIP_* */
  int with_normal;      /* 0/1 */
  int full_space;       /* 0 half space (default), 1 real-source term only */
  int store;            /* RSFG_PREC_* mask (FP64 always kept) */
  double scale;
  double nf_tol;        /* RSFG_PREC_BF16NF energy tolerance (<= 0: 2^-6) */
} rsfg_assembly;

int rsfg_assemble(rsfg_handle *h, const rsfg_assembly *a, double *gpu_ms);

/* operator dimensions after assembly: rows (shear), cols, leading dim */
int rsfg_operator_dims(rsfg_handle *h, int *rows, int *cols, int *ld);

/* bytes of the stored operator of precision prec that one product reads
   (all rows, shear and normal; RSFG_PREC_BF16NF includes its lo blocks).
   No reference counterpart: the algorithmic traffic of the product, for
   measurement. */
int rsfg_operator_bytes(rsfg_handle *h, int prec, double *bytes);

/* GenKEntries_petsc semantics (src/interact_matrix_assembly.c:102): the
   m x k block of the assembled FP64 shear operator with rows[] and cols[]
   (0-based), written column-major into out (out[r + m*c]).  With the
   normal rows assembled, rows ncomp*n .. ncomp*n + n - 1 address the "In"
   operator (row ncomp*n + p: normal traction at patch p, scaled by
   -scale, src/rsf_solve.c:245-247). */
int rsfg_entries(rsfg_handle *h, int m, const int *rows, int k,
                 const int *cols, double *out);

/* Unscaled traction blocks for arbitrary (receiver, source) pairs computed
   directly by the device kernel, internal units (shear modulus 1e4):
   out[6*p + 3*s + t] for slip s (strike, dip) and traction t (strike,
   dip, normal).  A bulk ckernel_func (src/eval_green.c:368-381) for
   CPU-side H-matrix builders. */
int rsfg_traction_pairs(rsfg_handle *h, long npairs, const int *recv,
                        const int *src, int full_space, double *out);

/* Vertical-edge tolerance [rad] of the triangles' free-surface correction
   for later rsfg_assemble and rsfg_traction_pairs calls.  An edge within
   the tolerance of vertical carries no correction term
   (AngSetupFSC_S, src/green/tdstresshs.f90:398-401, where the tolerance is
   EPS_FOR_FORTRAN, 5e-15; vertical_tol <= 0 restores it).  No reference
   counterpart as a setting: for edges a little further from vertical the
   correction's two end terms grow like cot(beta)^2 and cancel, and in FP64
   the reference's rule loses every digit of far entries there (on a rough
   32000-triangle mesh, errors up to 0.23 of a row's largest entry).
   3e-3 keeps such entries within 5e-7 of the row scale.  Must be < 0.1. */
int rsfg_set_triangle_tolerance(rsfg_handle *h, double vertical_tol);

/* ---- stress product ---------------------------------------------------
   MatMult(Is, x, y) (src/rsf_engine.c:168-170): y = K x with the stored
   operator of precision prec (one RSFG_PREC_* bit).  The reduced formats
   multiply in FP32 against an FP32 copy of x, sum each 16-byte slice of a
   row (4 or 8 entries) in FP32 and accumulate the slice sums in FP64; FP64
   is FP64 throughout.  x has cols entries, y rows entries. */
int rsfg_matvec(rsfg_handle *h, int prec, const double *x, double *y);
int rsfg_matvec_dev(rsfg_handle *h, int prec, const void *x_dev,
                    void *y_dev);
/* rsfg_matvec_dev without any allocation.  No counterpart in the
   reference: for callers that cannot call hipMalloc (code under HIP graph
   capture, tight serving loops, harnesses that forbid allocation).  The
   padded product input and, with normal rows, the product over all rows are
   staged in a static device array of 512 KiB per GPU (65536 doubles, held
   until the process ends) instead of the handle's allocated scratch, so the
   padded columns (cols + 1024 rounded up to a multiple of 512) plus, with
   normal rows, all rows must fit in 65536 doubles (e.g. 32000 DOF without
   normal rows); a larger operator returns RSFG_EINVAL.  Same
   kernels and results as rsfg_matvec_dev, asynchronous on the handle's
   stream.  Every call on a GPU uses that GPU's one array: calls that
   overlap on one GPU (other handles or threads) race. */
int rsfg_matvec_dev_noalloc(rsfg_handle *h, int prec, const void *x_dev,
                            void *y_dev);
/* transposed product K^T x (adjoint) */
int rsfg_matvec_t(rsfg_handle *h, int prec, const double *x, double *y);
/* time nrep back-to-back products on device-resident data; returns the
   mean kernel time [ms] measured with HIP events on the handle's stream */
int rsfg_matvec_bench(rsfg_handle *h, int prec, int nrep, double *ms);

/* Average time [ms] of nrep launches of a plain read stream over a fresh
   device buffer of `bytes` (1 MiB to 64 GiB) on the handle's device and
   stream: 16-byte nontemporal loads, eight in flight per lane, two
   workgroups of 8 waves per CU, nothing written.  No reference
   counterpart: the read rate the device reaches with the products' load
   pattern and no arithmetic, beside which their HBM rates can be read.
   Allocates and frees the buffer. */
int rsfg_read_bench(rsfg_handle *h, double bytes, int nrep, double *ms);

/* ---- rate-and-state problem -------------------------------------------
   Parameters of rsf_get_settings (src/rsf_init.c:273-503) that the RHS
   uses; per-patch a, b (read_rsf, src/interact/input.c:83) and D_c
   (-rsf_dc_file). */
typedef struct {
  int law;              /* 1 aging (default), 2 slip, 3 PRZ, 4 Sato, 5 KT */
  double v0, f0;        /* reference speed [m/s], friction */
  double eta;           /* radiation damping rd_fac G/(2 c_s) [Pa s/m] */
  double vmin_state;    /* |v| floor of laws 2,4,5 (1e-16) */
  double sato_beta;     /* 1e-2 */
  double kt_vc;         /* Kato-Tullis Vc, default 1e-2 v0 (<= 0: default) */
  int calc_sigma_dot;   /* evolve sigma with the In rows (-calc_sigma_dot) */
  int limit_sigma;      /* HBI limitsigma */
  double min_sigma, max_sigma;
} rsfg_friction;

int rsfg_set_friction(rsfg_handle *h, const rsfg_friction *f,
                      const double *a, const double *b, const double *dc);

/* backslip loading (src/rsf_solve.c:268-363): sinc = K (-vpl) with the
   FP64 operator, per-patch plate rate vpl_patch (NULL: uniform vpl)
   along the slip direction (strike component in the two-component
   mode), plus optional additive per-patch rates taudot_add,
   sigdot_add (-rsf_stress_rate_file; may be NULL). */
int rsfg_set_loading(rsfg_handle *h, double vpl, const double *vpl_patch,
                     const double *taudot_add, const double *sigdot_add);

/* rsf_ODE_RHSFunction (src/rsf_engine.c:186-302) with the operator of
   precision prec, and rsf_domain_check (:309-362) of X into *domain_ok.
   X, F are patch-major AoS (see conventions). */
int rsfg_rhs(rsfg_handle *h, int prec, const double *X, double *F,
             int *domain_ok);

/* ---- adaptive cycle integration ---------------------------------------
   PETSc TSRK (5dp/3bs) + TSAdapt basic + TSEvent + rsf_TS_Monitor as set
   up in src/rsf_solve.c:440-643, run entirely on the device: stage
   kernels stream K with the RHS fused into their epilogue, and a
   single-workgroup control kernel after each step attempt applies the
   step-size controller, locates log10 max|v| threshold crossings on the
   step's dense output, writes monitor and event rows, and advances the
   state.  The host replays captured graphs of step attempts and only
   polls a halted flag. */
typedef struct {
  int method;           /* RSFG_RK5DP / RSFG_RK3BS / RSFG_ARKIMEX_L2 */
  int prec;             /* operator precision used in the stages */
  double rtol;          /* scalar relative tolerance (-rtol) */
  double atol[6];       /* per state component (psi, tau, ...) */
  double dt_init, dt_max;           /* [s] */
  double dt_monitor;    /* [s] monitor cadence, also caps the step */
  double rdx_monitor, adx_monitor;
  double vel_event, vel_event_hyst, rupture_vth;
  long max_steps;
  int max_domain_reject;            /* storm guard (<= 0 off) */
  int steps_per_launch;             /* accepted-step attempts per graph */
  int max_events;       /* rows of the event log (<= 0: 4096) */
  int max_monitor;      /* rows of the monitor log (<= 0: 262144) */
} rsfg_solver_opts;

/* fills o with the reference defaults (src/rsf_init.c:331-350) */
int rsfg_solver_defaults(rsfg_solver_opts *o);

/* sets the state (AoS, reference layout) at time t */
int rsfg_set_state(rsfg_handle *h, double t, const double *X);
int rsfg_get_state(rsfg_handle *h, double *t, double *X);

/* ---- input tables and checkpoints ------------------------------------
   Per-patch input tables of rsf_solve: n rows of ncols whitespace-separated
   numbers, row-major into out: the rsf file "a b" (read_rsf,
   src/interact/input.c:83-103), -rsf_ic_file "tau v", -rsf_dc_file "dc",
   -rsf_sigma_file "sigma0" (src/rsf_solve.c:162-218), -rsf_vpl_file "vpl"
   and -rsf_stress_rate_file "tau_dot sigma_dot" (:275-360).  A short or
   malformed file returns RSFG_EINVAL naming the row (the reference exits). */
int rsfg_read_table(const char *fname, int n, int ncols, double *out);

/* rsf_solve checkpoint (rsf_write_checkpoint / rsf_read_checkpoint,
   src/rsf_output.c:595-716): a PETSc binary file (big endian) holding a
   16-entry metadata vector (magic 20260726, version 1, patches, state dim,
   slip mode, state law, step, t, dt) and the state vector in the reference's
   patch-major layout.  The writer goes through <file>.tmp and a rename and
   keeps the previous file as <file>.prev.  The reader checks the magic, the
   patch count and the state dim, and loads the state into the handle (as
   rsfg_set_state); slip_mode, law, step and dt may be NULL. */
int rsfg_write_checkpoint(rsfg_handle *h, const char *fname, int slip_mode, int law,
                          long step, double dt);
int rsfg_read_checkpoint(rsfg_handle *h, const char *fname, int *slip_mode, int *law,
                         long *step, double *t, double *dt);

typedef struct {
  long steps, rejects, domain_rejects, rhs_evals, launches;
  double t, dt;         /* time reached and next step size [s] */
  int status;           /* 0 ok, RSFG_ESTORM, RSFG_EDOMAIN */
  int nevent, nmonitor; /* rows available */
  double wall_s;        /* host wall time of the solve call */
  int events_dropped;   /* event / monitor rows past the logs' capacity */
  int monitor_dropped;  /*   (max_events, max_monitor): counted, not stored */
  int ncatalog;         /* catalog rows available */
  int catalog_dropped;  /* catalog rows past the capacity (max_events) */
} rsfg_stats;

/* integrates from the current state to t_end (STEPOVER semantics) */
int rsfg_solve(rsfg_handle *h, const rsfg_solver_opts *o, double t_end,
               rsfg_stats *st);
/* continues the last rsfg_solve (or rsfg_solve_continue) of this handle to
   a later t_end.  It resumes from the true state at the last accepted step
   (STEPOVER, src/rsf_solve.c:506), with the next step size kept as a
   checkpoint restart keeps it (:634-640), never from an interpolated
   state (the drift the reference warns of at :622-632): the step size, the
   FSAL derivative, the event state (slipping flag, onset snapshots) and the
   monitor reference are kept, event, monitor and step-log rows are
   appended, and the accepted steps are exactly those of one rsfg_solve to
   the final t_end.  o->method and o->prec must be those of the solve being
   continued (o->dt_init is unused); any call that changes or reuses the
   solver buffers in between (rsfg_set_state, rsfg_set_friction,
   rsfg_set_loading, rsfg_rhs, rsfg_stage_bench, the adjoint calls, a
   checkpoint read) returns RSFG_ESTATE here.  Statistics count from the
   first solve. */
int rsfg_solve_continue(rsfg_handle *h, const rsfg_solver_opts *o, double t_end,
                        rsfg_stats *st);

/* ---- ensembles ---------------------------------------------------------
   Batched runs over shared operators (no counterpart in the reference,
   which runs one parameter set per process): members_per_geom members per
   geometry handle (16, 32 or 64; up to 64 geometry handles), each a member
   handle (rsfg_create_member)
   configured through the ordinary calls (rsfg_set_friction,
   rsfg_set_loading, rsfg_set_state; rsfg_events, rsfg_monitor,
   rsfg_get_state afterwards).  Every geometry needs an FP32 operator.  The
   solve advances all members together, each with its own adaptive steps,
   events and monitor; each stage of a geometry's members is one product
   K [v_1 .. v_M] on the matrix cores (FP32 operands and accumulation,
   flushed to FP64 every 256 columns).  Member m of geometry g is
   g members_per_geom + k. */
typedef struct rsfg_ensemble rsfg_ensemble;
int rsfg_ensemble_create(rsfg_handle *const *geoms, int ngeom, int members_per_geom,
                         rsfg_ensemble **e);
int rsfg_ensemble_member(rsfg_ensemble *e, int member, rsfg_handle **h);
/* operator of the batched product: RSFG_PREC_FP32 (the default) or
   RSFG_PREC_BF16NF, a two-byte copy built from each geometry's FP32
   operator: bf16 hi of every entry in tiles for v_mfma_f32_16x16x32_bf16,
   plus the bf16 lo (the FP32 entry minus hi) of every 256-row group over
   the hull of the 256-column segments its rows need: per row, the segments
   in decreasing energy until those left out hold at most nf_tol^2 of the
   row's energy (nf_tol <= 0: 2^-6, the single-fault RSFG_PREC_BF16NF
   default).
   The member slip rates enter as three bf16 terms whose sum is their FP32
   value, so each tile's product is its bf16 entries times the FP32 slip
   rates, accumulated in FP32 and flushed to FP64 every 256 columns.
   RSFG_PREC_BF16NF takes 16 or 32 members per geometry and runs on the
   stream-K shapes; switching keeps the shape when it fits, else takes the
   first that does.  bytes (may be NULL): operator bytes one product
   streams (hi, lo and padding).  No counterpart in the reference. */
int rsfg_ensemble_set_operator(rsfg_ensemble *e, int prec, double nf_tol, double *bytes);
/* o->prec must name the ensemble's operator (RSFG_PREC_FP32 unless
   rsfg_ensemble_set_operator chose RSFG_PREC_BF16NF), method rk5dp or
   rk3bs; st: one record per member */
int rsfg_ensemble_solve(rsfg_ensemble *e, const rsfg_solver_opts *o, double t_end,
                        rsfg_stats *st);
/* autotuner: times every candidate tile shape of the batched stage on this
   ensemble (members must be configured), keeps the fastest; ms (may be
   NULL) receives one time per candidate, -1 where a shape does not fit */
int rsfg_ensemble_tune(rsfg_ensemble *e, const rsfg_solver_opts *o, int *choice, double *ms);
/* current unit shape, number of candidates, and per candidate seven ints
   (operator rows per wave, waves per product unit, 16-column chunks per
   load group, columns per product unit, load groups each wave keeps in its
   register ring, 1 where the unit's waves share the member slip rates
   through LDS, and for stream-K shapes the workgroups per CU, each taking
   an equal share of the product, with 0 columns per unit; 0 otherwise) */
int rsfg_ensemble_config(rsfg_ensemble *e, int *cfg, int *ncfg, int *shape);
/* use candidate shape cfg (as listed by rsfg_ensemble_config) without
   tuning, e.g. a shape tuned earlier on the same layout; RSFG_EINVAL when
   it is out of range or does not fit the members per geometry */
int rsfg_ensemble_select(rsfg_ensemble *e, int cfg);
/* mean durations by HIP events of nrep launches each of the current
   shape's product kernel and stage epilogue (a middle stage at the
   members' states); measurement only */
int rsfg_ensemble_bench(rsfg_ensemble *e, const rsfg_solver_opts *o, int nrep,
                        double *ms_product, double *ms_epilogue);
int rsfg_ensemble_destroy(rsfg_ensemble *e);

/* times nrep back-to-back launches of the fused stage kernel (K v with the
   RHS epilogue, as used inside rsfg_solve) at the current state with the
   operator of precision prec; *ms is the mean kernel duration measured with
   HIP events on the handle's stream.  No reference counterpart: the
   reference's per-stage cost is MatMult + rsf_ODE_RHSFunction. */
int rsfg_stage_bench(rsfg_handle *h, int prec, int nrep, double *ms);

/* event rows (rsf_events.dat + catalog, src/rsf_output.c:1000-1150):
   8 doubles each: t [s], +1 onset / -1 arrest, log10 max|v|, mean slip,
   mean mu, peak |v| since onset, mean stress drop over ruptured cells
   [Pa], ruptured cell count.  *n in: capacity, out: rows copied. */
int rsfg_events(rsfg_handle *h, double *rows, int *n);
/* catalog rows (rsf_catalog.dat, rsf_finalize_event, src/rsf_output.c:
   1078-1180), one per closed event that ruptured at least one cell (|v| >=
   rupture_vth at an accepted step between its onset and arrest): 11 doubles
   each: onset t [s], arrest t [s], ruptured cells, their area (4 l w per
   cell; a triangle counts 4 times its area, as in the reference, whose
   triangles carry l = w = sqrt(area)), mean and max |slip change| [m],
   mean and max stress drop |tau_onset| - |tau_arrest| [Pa], peak |v|,
   M0 = G sum area |slip change| [N m] with G = scale 1e4, and
   Mw = 2/3 (log10 M0 - 9.05) (-99 for M0 <= 0).  In the two-component
   layout tau and slip are the strike components, as in the event rows.
   *n in: capacity, out: rows copied. */
int rsfg_catalog(rsfg_handle *h, double *rows, int *n);
/* monitor rows (rsf_monitor.dat, src/rsf_output.c:832-842): 10 doubles
   each: step, t [s], t [yr], dt, log10 max|v|, mean slip, mean mu,
   max sigma, min sigma, log10 mean|v| */
int rsfg_monitor(rsfg_handle *h, double *rows, int *n);

/* ---- sensitivities: discrete adjoint of the explicit RK integration ----
   No counterpart in the reference, which has no adjoint (its nearest code
   is the analytic per-cell Jacobian of the IMEX path, src/rsf_imex.c:165-237).
   The differentiated map is a replay of rsf_ODE_RHSFunction
   (src/rsf_engine.c:186-302) under the rk5dp / rk3bs tableau over a fixed
   list of step sizes, normally the accepted steps of an rsfg_solve
   (rsfg_set_step_log); the objective is the misfit of log slip rates at
   receiver patches,
     J = 1/2 sum_n weight[n] sum_r (ln|v_{recv[r]}(X_n)| - target[n][r])^2,
   over the states X_0 (the handle's current state) .. X_nsteps.  Operators
   with normal-stress rows are not supported. */

/* keep the accepted step sizes of subsequent solves (capacity steps; 0
   turns the log off) */
int rsfg_set_step_log(rsfg_handle *h, long long capacity);
/* the accepted step sizes of the last solve: *nsteps = their number, the
   first min(*nsteps, capacity) are copied to steps (may be NULL) */
int rsfg_get_step_log(rsfg_handle *h, double *steps, long long capacity, long long *nsteps);

typedef struct {
  int method;            /* RSFG_RK5DP or RSFG_RK3BS */
  int prec;              /* operator of K v and K^T w: RSFG_PREC_FP64 or RSFG_PREC_FP32 */
  int checkpoint_every;  /* steps between stored states of the adjoint (<= 0: 256) */
  int nrecv;             /* receiver patches */
  const int *recv;       /* [nrecv] patch indices */
  const double *target;  /* [(nsteps + 1) x nrecv] target ln|v|, or NULL (replay only) */
  const double *weight;  /* [nsteps + 1] weight of each state, or NULL (replay only) */
} rsfg_adjoint_problem;

/* replays nsteps steps from the handle's current state (which is left
   unchanged).  J (if not NULL) the objective; trace (if not NULL)
   [(nsteps + 1) x nrecv] ln|v| at the receivers; X_end (if not NULL) the
   final state in rsfg_set_state's layout */
int rsfg_replay(rsfg_handle *h, const double *steps, long long nsteps,
                const rsfg_adjoint_problem *pb, double *J, double *trace, double *X_end);
/* J and its gradient with respect to the per-patch friction coefficients
   a, b and D_c ([n] each, any may be NULL): forward replay with
   checkpoints every checkpoint_every steps, then the backward sweep over
   the steps with each segment recomputed from its checkpoint */
int rsfg_adjoint_gradient(rsfg_handle *h, const double *steps, long long nsteps,
                          const rsfg_adjoint_problem *pb, double *J, double *grad_a,
                          double *grad_b, double *grad_dc);
/* times nrep launches of the backward sweep's transposed product K^T w
   (row-split partial sums, FP64 or FP32 operator): *ms is the mean kernel
   duration by HIP events on the handle's stream */
int rsfg_adjoint_bench(rsfg_handle *h, int prec, int nrep, double *ms);

#ifdef __cplusplus
}
#endif
#endif
