/*
  rsfo: FP64 CPU restatement of the reference rsf_solve dense hot path.

  TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the CPU
  baseline; the product (interact_amd, librsfg) never links or calls it.

  Every function names the reference file:line whose behaviour it restates.
*/
#ifndef RSFO_H
#define RSFO_H
#ifdef __cplusplus
extern "C" {
#endif

/* Okada (1992) DC3D for a finite rectangle, strike (d1) and dip (d2) slip.
   Restates src/green/dc3d.F:576-813 (DC3D), :1161-1196 (DCCON0),
   :1254-1320 (DCCON2), :816-911 (UA), :913-1040 (UB), :1042-1159 (UC).
   Tensile opening (DISL3) is not part of the operator and not restated.
   u[12] = ux uy uz uxx uyx uzx uxy uyy uzy uxz uyz uzz; returns iret. */
int rsfo_dc3d(double alpha, double x, double y, double z, double depth,
              double dip_deg, double al1, double al2, double aw1, double aw2,
              double d1, double d2, int full_space, double u[12]);

/* geometry preprocessing of one rectangular patch, as read_geometry does
   it (src/read_geometry.c:412-447; strike, dip are float in struct flt,
   src/includes/structures.h:69).  frame[13] = cos_alpha sin_alpha
   normal[3] t_strike[3] t_dip[3] dip_deg(double of the float) spare */
void rsfo_patch_frame(float strike, float dip, double *frame);

/* tractions (strike, dip, normal) at every receiver centroid for unit
   strike and unit dip slip on every source, internal units
   (shear modulus 1e4, nu 0.25; src/read_geometry.c:53, 702-719).
   Restates GenKEntries_petsc (src/interact_matrix_assembly.c:102-151) ->
   eval_green_at_receiver (src/eval_green.c:244) -> eval_okada
   (src/green/eval_okada.c:59-175) -> resolve_stress_on_fault
   (src/geometry/geometry.c:84-111).  Singular pairs give 0.
   T is double[n][n][2][3] (receiver-major); iret may be NULL.
   pairs are split over nthreads OpenMP threads. */
void rsfo_traction_tensor(int n, const double *xyz, const float *strike,
                          const float *dip, const double *l, const double *w,
                          int full_space, int nthreads, double *T, int *iret);

/* single pair, same semantics (used by the op counter and tests) */
void rsfo_traction_pair(const double *xr, const double *fr_rec,
                        const double *xs, const double *fr_src, double ls,
                        double ws, int full_space, double out[2][3],
                        int iret[2]);

/* ---- triangular dislocations (Nikkhoo & Walter 2015), rsfo_tri.c ---- */

/* TDCS base vectors and area of the triangle p1 p2 p3
   (get_tdcs_base_vectors, src/green/tddisphs.f90:790-817) */
void rsfo_tri_frame(const double *p1, const double *p2, const double *p3,
                    double *vnorm, double *vstrike, double *vdip, double *area);

/* stress S at point P (z <= 0) from slip (strike ss, dip ds, tensile ts)
   on the triangle p1 p2 p3 in a half space (tdstresshs,
   src/green/tdstresshs.f90:9-148) with Poisson ratio nu, 2 mu and lambda.
   Returns 0, 1 if the stress is not finite (P on an edge line in the
   triangle's plane: the reference's NaN -> entry 0), -1 if a point lies
   above the surface.  A triangle with all vertices at z = 0 gives S = 0. */
int rsfo_tri_stress(const double *P, const double *p1, const double *p2,
                    const double *p3, double ss, double ds, double ts,
                    double nu, double mu2, double lambda, double S[3][3]);

/* vertical-edge tolerance [rad] of the free-surface correction: an edge
   whose angle from the downward vertical is within it of 0 or pi carries
   no correction (AngSetupFSC_S, src/green/tdstresshs.f90:398-401, with
   EPS_FOR_FORTRAN = 5e-15, restored by tol <= 0).  Process-wide; not
   for use while an evaluation runs. */
void rsfo_set_tri_vertical_tol(double tol);

/* tractions (strike, dip, normal) at every receiver for unit strike and
   unit dip slip on every source of a mesh of rectangles and triangles,
   internal units (shear modulus 1e4, nu 0.25), as GenKEntries_petsc sees
   them.  Per patch: is_tri, receiver point xyz[3] (the centroid),
   fr[13] in rsfo_patch_frame's layout (for triangles only the normal,
   strike and dip vectors, fr[2..10], are used: the TDCS vectors), l, w
   (rectangles), verts[9] (triangles).  T and iret as rsfo_traction_tensor;
   triangle sources are always evaluated in the half space. */
void rsfo_mesh_traction_tensor(int n, const int *is_tri, const double *xyz,
                               const double *fr, const double *l,
                               const double *w, const double *verts,
                               int full_space, int nthreads, double *T,
                               int *iret);

/* the same tractions for np (recv[p], source src[p]) pairs of the mesh
   into T[6 p + 3 s + t] */
void rsfo_mesh_traction_pairs(int n, const int *is_tri, const double *xyz,
                              const double *fr, const double *l, const double *w,
                              const double *verts, int full_space, int nthreads,
                              long long np, const int *recv, const int *src,
                              double *T);

/* patch-file reader with the reference's scan format
   "%lf %lf %lf %f %f %lf %lf %i" (src/includes/precision_double.h:31,
   read loop src/read_geometry.c:115-117).  Rectangles only.  Returns the
   number of patches read (at most nmax), or -1 if the file cannot be read.
   xyz[3n], strike[n], dip[n] (float), l[n], w[n], group[n] */
int rsfo_read_patches(const char *fname, int nmax, double *xyz, float *strike,
                      float *dip, double *l, double *w, int *group);

/* ---- rate-and-state physics (src/rsf_engine.c) ---- */
typedef struct {
  int n;                 /* patches */
  int ncomp;             /* 1 = reference scalar slip, 2 = vector slip */
  const double *K;       /* (ncomp n) x (ncomp n) row-major shear operator, SI */
  const double *Kn;      /* n x (ncomp n) normal operator (compression +), or NULL */
  const double *a, *b, *dc;      /* per patch */
  const double *sinc;    /* per row of K: backslip shear loading rate [Pa/s] */
  const double *sinc_n;  /* per patch: normal loading rate, or NULL */
  double v0, f0, eta, vmin_state, sato_beta, kt_vc;
  int law;               /* 1 aging 2 slip 3 PRZ 4 Sato 5 Kato-Tullis */
  int calc_sigma_dot, limit_sigma;
  double min_sigma, max_sigma;
  int nthreads;
} rsfo_problem;

/* vel_from_rsf, src/rsf_engine.c:20-30 */
double rsfo_vel(double tau, double sigma, double psi, double a, double v0);
/* rsf_state_rate, src/rsf_engine.c:44-132 (dSdpsi, dSdv may be NULL) */
double rsfo_state_rate(const rsfo_problem *p, int i, double psi, double v,
                       double *dSdpsi, double *dSdv);
/* rsf_ODE_RHSFunction, src/rsf_engine.c:186-302, X and F AoS:
   scalar: (psi, tau, sigma, slip) per patch;
   vector: (psi, tau_s, tau_d, sigma, slip_s, slip_d) per patch.
   vel (ncomp n) and taudot (ncomp n) may be NULL */
void rsfo_rhs(const rsfo_problem *p, const double *X, double *F,
              double *vel, double *taudot);
/* rsf_domain_check, src/rsf_engine.c:309-362; returns 1 if acceptable */
int rsfo_domain_ok(const rsfo_problem *p, const double *X);

/* ---- IMEX split (src/rsf_imex.c), scalar layout only ---- */
/* explicit part G (rsf_IMEX_RHSFunction) and implicit part Fimpl
   (rsf_IMEX_IFunction as Xdot - Fimpl), AoS like rsfo_rhs; G + Fimpl = F */
void rsfo_imex_parts(const rsfo_problem *p, const double *X, double *G, double *Fi);
/* dFimpl/dX of cell i at x (4x4 row-major); rsf_IMEX_IJacobian is
   shift I minus this block */
void rsfo_imex_block(const rsfo_problem *p, int i, const double *x, double *J);
/* stage equation Y = Z + hg Fimpl(Y) of cell i by per-cell Newton;
   y: initial guess (psi, tau) in, stage state (4) out; fi: (Y - Z)/hg.
   Returns the iteration count or -1 */
int rsfo_imex_newton(const rsfo_problem *p, int i, const double *Z, double hg,
                     double rtol, const double *atol, double *y, double *fi);

/* ---- adaptive integrator, events and monitor ---- */
typedef struct {
  int method;                 /* 0 = rk5dp (default), 1 = rk3bs,
                                 2 = ARKIMEX l2 (scalar layout) */
  double rtol;
  double atol[6];             /* per state component (dim of the layout) */
  double t0, t_end, dt_init, dt_max;
  double dt_monitor, rdx_monitor, adx_monitor;
  double vel_event, vel_event_hyst, rupture_vth;
  long max_steps;
  int max_domain_reject;      /* storm guard, <= 0 disables */
  int max_events, max_monitor;
} rsfo_options;

typedef struct {
  long steps, rejects, domain_rejects, rhs_evals;
  double t_final, dt_final;
  int status;                 /* 0 ok, 1 storm guard, 2 max steps, 3 nonfinite */
  int nevent;                 /* rows written to ev */
  double *ev;                 /* max_events x 8: t, +1/-1, log10 max|v|,
                                 mean slip, mean mu, peak |v| since onset,
                                 mean stress drop over ruptured cells [Pa],
                                 ruptured cell count */
  int nmon;
  double *mon;                /* max_monitor x 10 monitor columns */
  double *hlog;               /* accepted step sizes in order, or NULL */
  long hlog_cap;
  /* catalog (rsf_finalize_event, src/rsf_output.c:1078-1180), or cat NULL:
     one row of 11 per closed event with ruptured cells, at most max_events:
     onset t, arrest t, ruptured cells, their area, mean and max |slip
     change|, mean and max stress drop, peak |v|, M0 = cat_mu sum area
     |slip change|, Mw; cat_area: per-patch area 4 l w (NULL: zero) */
  int ncat;
  double *cat;
  const double *cat_area;
  double cat_mu;
} rsfo_result;

/* integrates X in place; X is AoS as for rsfo_rhs */
int rsfo_solve(const rsfo_problem *p, const rsfo_options *o, double *X,
               rsfo_result *r);

/* one adaptive step attempt from (t, X) with step h: writes the new state,
   the weighted error norm and the stage domain flag (tests) */
int rsfo_rk_attempt(const rsfo_problem *p, int method, double rtol,
                    const double *atol, double h, const double *X,
                    const double *F0, double *Xnew, double *Fnew,
                    double *errnorm);

/* ---- discrete adjoint of the replayed integration, rsfo_adj.c ----
   (no reference counterpart; the map is rsfo_rk_attempt stepping over a
   fixed step list, the objective
   J = 1/2 sum_n weight[n] sum_r (ln|v_recv[r](X_n)| - target[n][r])^2) */

/* right-hand side f[dim] and slip rate v[nc] of patch i from its state x and
   stressing rates td[nc] (complex restatement evaluated on the real axis) */
void rsfo_patch_rhs_c(const rsfo_problem *p, int i, const double *x, const double *td,
                      double *f, double *v);
/* complex-step Jacobians of patch i (row-major, any output may be NULL):
   Jx[dim][dim] df/dx, Jtd[dim][nc] df/dtd, Jp[dim][3] df/d(a, b, dc),
   Vx[nc][dim] dv/dx, Vp[nc] dv/da, Lx[dim] dln|v|/dx, *La dln|v|/da */
void rsfo_patch_jacobian(const rsfo_problem *p, int i, const double *x, const double *td,
                         double *Jx, double *Jtd, double *Jp, double *Vx, double *Vp,
                         double *Lx, double *La);
/* replay nsteps steps (method 0 rk5dp, 1 rk3bs) from X0 (AoS); trace
   [(nsteps+1) x nrecv] ln|v| at the receivers and Xend may be NULL */
void rsfo_replay(const rsfo_problem *p, int method, const double *X0, const double *steps,
                 long nsteps, const int *recv, int nrecv, double *trace, double *Xend);
/* J (returned) and dJ/da, dJ/db, dJ/ddc per patch; stores every stage */
double rsfo_adjoint(const rsfo_problem *p, int method, const double *X0, const double *steps,
                    long nsteps, const int *recv, int nrecv, const double *target,
                    const double *weight, double *ga, double *gb, double *gdc);

#ifdef __cplusplus
}
#endif
#endif
