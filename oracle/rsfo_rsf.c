/*
  Oracle restatement of the rsf_solve right-hand side, domain check and the
  PETSc TSRK + TSAdapt(basic) + TSEvent + monitor loop around it.

  TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).

  The reference integrates with PETSc 3.19.6 (third party, not vendored):
  rk5dp / rk3bs tableaus, infinity-norm weighted error, basic controller
  (safety 0.9, 0.5 extra after a failed attempt, exponent -1/order, clip
  [0.5, 2], h <= min(dt_max, dt_monitor)), stage domain failures scale h by
  0.25 (src/rsf_solve.c:440-546).  Events are located on the step's dense
  output instead of by re-stepping (the reference lands steps on events);
  both locate the log10 max|v| crossing, here to 1e-10 in log10 v.
*/
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "rsfo.h"

#define SEC_PER_YEAR 31557600.0               /* src/includes/constants.h:31 */
#define ARG_MAX 600.0                          /* src/rsf_engine.c:316 */

/* vel_from_rsf, src/rsf_engine.c:20-30 */
double rsfo_vel(double tau, double sigma, double psi, double a, double v0)
{
  double mu = tau / sigma;
  double st = mu / a;
  double ef = exp(-psi / a);
  return 2. * v0 * ef * sinh(st);
}

/* rsf_state_rate, src/rsf_engine.c:44-132 */
double rsfo_state_rate(const rsfo_problem *p, int i, double psi, double vin,
                       double *dSdpsi, double *dSdv)
{
  double S = 0.0, vabs, b = p->b[i], D = p->dc[i], e, om, gate, lom, pss;
  int floored = 0;
  if (b == 0.0) {
    if (dSdpsi) *dSdpsi = 0.0;
    if (dSdv) *dSdv = 0.0;
    return 0.0;
  }
  switch (p->law) {
  case 1: /* aging */
    e = exp((p->f0 - psi) / b);
    S = (b / D) * (p->v0 * e - fabs(vin));
    if (dSdpsi) *dSdpsi = -(p->v0 / D) * e;
    if (dSdv) *dSdv = -b / D;
    break;
  case 2: /* slip */
    vabs = fabs(vin);
    if (vabs < p->vmin_state) { vabs = p->vmin_state; floored = 1; }
    pss = p->f0 - b * log(vabs / p->v0);
    S = -(vabs / D) * (psi - pss);
    if (dSdpsi) *dSdpsi = -vabs / D;
    if (dSdv) *dSdv = floored ? 0.0 : (-(psi - pss) / D - b / D);
    break;
  case 3: /* PRZ */
    vabs = fabs(vin);
    e = exp((psi - p->f0) / b);
    S = (b / (2.0 * D)) * (p->v0 / e - (vabs * vabs / p->v0) * e);
    if (dSdpsi) *dSdpsi = -(1.0 / (2.0 * D)) * (p->v0 / e + (vabs * vabs / p->v0) * e);
    if (dSdv) *dSdv = -(b / D) * (vabs / p->v0) * e;
    break;
  case 4: /* Sato */
  case 5: /* Kato-Tullis */
    vabs = fabs(vin);
    if (vabs < p->vmin_state) { vabs = p->vmin_state; floored = 1; }
    e = exp((psi - p->f0) / b);
    om = (vabs / p->v0) * e;
    gate = (p->law == 4) ? exp(-om / p->sato_beta) : exp(-vabs / p->kt_vc);
    lom = log(vabs / p->v0) + (psi - p->f0) / b;
    S = (b / D) * ((p->v0 / e) * gate - vabs * lom);
    if (dSdpsi)
      *dSdpsi = (b / D) * ((p->v0 / e) * (((p->law == 4) ? (-gate * om / (b * p->sato_beta)) : 0.0) - gate / b) - vabs / b);
    if (dSdv) {
      if (floored) *dSdv = 0.0;
      else *dSdv = (b / D) * ((p->v0 / e) * ((p->law == 4) ? (-gate * e / (p->v0 * p->sato_beta)) : (-gate / p->kt_vc)) - lom - 1.0);
    }
    break;
  default:
    S = NAN;
  }
  return S;
}

static int state_dim(const rsfo_problem *p) { return p->ncomp == 2 ? 6 : 4; }

/* slip-rate vector of one patch from its state */
static void patch_vel(const rsfo_problem *p, int i, const double *x, double *v)
{
  if (p->ncomp == 1) {
    v[0] = rsfo_vel(x[1], x[2], x[0], p->a[i], p->v0);
  } else {
    double tm = sqrt(x[1] * x[1] + x[2] * x[2]);
    double vm = 2.0 * p->v0 * exp(-x[0] / p->a[i]) * sinh(tm / (x[3] * p->a[i]));
    if (tm > 0.0) {
      v[0] = vm * x[1] / tm;
      v[1] = vm * x[2] / tm;
    } else {
      v[0] = v[1] = 0.0;
    }
  }
}

static void dense_matvec(int rows, int cols, const double *K, const double *x,
                         double *y, int nthreads)
{
  int i;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (i = 0; i < rows; i++) {
    const double *kr = K + (size_t)i * cols;
    double s = 0.0;
    int j;
    for (j = 0; j < cols; j++) s += kr[j] * x[j];
    y[i] = s;
  }
}

/* rsf_ODE_RHSFunction (src/rsf_engine.c:186-302) incl.
   rsf_compute_vel_and_stressing (:139-172) */
void rsfo_rhs(const rsfo_problem *p, const double *X, double *F, double *vel_out,
              double *taudot_out)
{
  int n = p->n, nc = p->ncomp, dim = state_dim(p), i;
  int nd = nc * n;
  double *vel = (double *)malloc(sizeof(double) * nd);
  double *td = (double *)malloc(sizeof(double) * nd);
  double *sd = p->calc_sigma_dot ? (double *)malloc(sizeof(double) * n) : NULL;
  for (i = 0; i < n; i++) patch_vel(p, i, X + (size_t)i * dim, vel + nc * i);
  dense_matvec(nd, nd, p->K, vel, td, p->nthreads);
  if (p->calc_sigma_dot) dense_matvec(n, nd, p->Kn, vel, sd, p->nthreads);
  for (i = 0; i < n; i++) {
    const double *x = X + (size_t)i * dim;
    double *f = F + (size_t)i * dim;
    double a = p->a[i], eta = p->eta, sdot, sig = x[nc + 1];
    double vmag = (nc == 1) ? vel[i] : sqrt(vel[2 * i] * vel[2 * i] + vel[2 * i + 1] * vel[2 * i + 1]);
    /* state evolution first (:262), with the signed scalar velocity in the
       reference layout; laws only use |v| */
    f[0] = rsfo_state_rate(p, i, x[0], (nc == 1) ? vel[i] : vmag, NULL, NULL);
    /* normal stress rate (:266-275) */
    sdot = p->calc_sigma_dot ? sd[i] + (p->sinc_n ? p->sinc_n[i] : 0.0) : 0.0;
    if (p->limit_sigma && ((sig < p->min_sigma) || (sig > p->max_sigma))) sdot = 0.0;
    f[nc + 1] = sdot;
    if (nc == 1) {
      /* shear stress rate from the quasi-dynamic balance (:280-290) */
      double mu = x[1] / x[2];
      double st = mu / a;
      double ef = exp(-x[0] / a);
      double pre = 2.0 * p->v0 / (a * x[2]);
      double cf = cosh(st) * ef;
      double dvdtau = pre * cf;
      double dvdsig = -pre * cf * mu;
      double dvdpsi = -vel[i] / a;
      f[1] = (td[i] + p->sinc[i] - eta * (dvdsig * f[2] + dvdpsi * f[0]));
      f[1] /= (1.0 + eta * dvdtau);
      f[3] = vel[i];
    } else {
      /* two slip components: the same balance, solved with the 2x2
         Jacobian dV/dtau = g_par t t^T + g_perp (I - t t^T), t = tau/|tau| */
      double ts = x[1], tdp = x[2], tm = sqrt(ts * ts + tdp * tdp);
      double arg = tm / (sig * a), ef = exp(-x[0] / a);
      double gpar = 2.0 * p->v0 / (a * sig) * cosh(arg) * ef;
      double gperp = (arg > 1e-8) ? vmag / tm : 2.0 * p->v0 / (a * sig) * ef;
      double th0 = (tm > 0.0) ? ts / tm : 1.0, th1 = (tm > 0.0) ? tdp / tm : 0.0;
      double dvmdsig = -gpar * tm / sig;          /* d|V|/dsigma */
      double r0, r1, rp;
      r0 = td[2 * i] + p->sinc[2 * i] - eta * (dvmdsig * th0 * f[3] - vel[2 * i] / a * f[0]);
      r1 = td[2 * i + 1] + p->sinc[2 * i + 1] - eta * (dvmdsig * th1 * f[3] - vel[2 * i + 1] / a * f[0]);
      rp = th0 * r0 + th1 * r1;
      f[1] = th0 * rp / (1.0 + eta * gpar) + (r0 - th0 * rp) / (1.0 + eta * gperp);
      f[2] = th1 * rp / (1.0 + eta * gpar) + (r1 - th1 * rp) / (1.0 + eta * gperp);
      f[4] = vel[2 * i];
      f[5] = vel[2 * i + 1];
    }
  }
  if (vel_out) memcpy(vel_out, vel, sizeof(double) * nd);
  if (taudot_out) memcpy(taudot_out, td, sizeof(double) * nd);
  free(vel);
  free(td);
  free(sd);
}

/* ---------------- IMEX split (src/rsf_imex.c, scalar layout) ----------------
   Per cell x = (psi, tau, sigma, slip), eta = G/(2 cs), denom = 1 + eta dv/dtau.
   Explicit G (rsf_IMEX_RHSFunction, src/rsf_imex.c:62-115):
     G_psi = 0, G_sigma = sigma_dot, G_tau = (tau_dot + sinc - eta dv/dsigma G_sigma) / denom,
     G_slip = v.
   Implicit Fimpl (rsf_IMEX_IFunction :123-155 written as F = Xdot - Fimpl):
     Fimpl_psi = S(psi, |v|), Fimpl_tau = -eta (dv/dpsi) S / denom, others 0.
   G + Fimpl equals rsf_ODE_RHSFunction. */
void rsfo_imex_parts(const rsfo_problem *p, const double *X, double *G, double *Fi)
{
  int n = p->n, i;
  double *F = (double *)malloc(sizeof(double) * 4 * n);
  double *vel = (double *)malloc(sizeof(double) * n);
  double *td = (double *)malloc(sizeof(double) * n);
  rsfo_rhs(p, X, F, vel, td);
  for (i = 0; i < n; i++) {
    const double *x = X + 4 * (size_t)i;
    double *g = G + 4 * (size_t)i, *fi = Fi + 4 * (size_t)i;
    double a = p->a[i], eta = p->eta, sdot = F[4 * (size_t)i + 2];
    double mu = x[1] / x[2], st = mu / a, ef = exp(-x[0] / a);
    double pre = 2.0 * p->v0 / (a * x[2]), cf = cosh(st) * ef;
    double dvdtau = pre * cf, dvdsig = -pre * cf * mu, denom = 1.0 + eta * dvdtau;
    double S = rsfo_state_rate(p, i, x[0], vel[i], NULL, NULL);
    g[0] = 0.0;
    g[2] = sdot;
    g[1] = (td[i] + p->sinc[i] - eta * (dvdsig * sdot)) / denom;
    g[3] = vel[i];
    fi[0] = S;
    fi[1] = -(eta * (-vel[i] / a) * S) / denom;
    fi[2] = 0.0;
    fi[3] = 0.0;
  }
  free(F); free(vel); free(td);
}

/* local implicit part of cell i at (psi, tau; sigma): Fimpl_psi, Fimpl_tau
   and, if J != NULL, the 4x4 block dFimpl/dX (row-major, rows psi and tau
   with columns psi, tau, sigma; the reference's IJacobian,
   src/rsf_imex.c:162-237, is shift I minus this block) */
static void imex_local(const rsfo_problem *p, int i, const double *x, double *fpsi,
                       double *ftau, double *J)
{
  double a = p->a[i], E = p->eta, sig = x[2];
  double mu = x[1] / sig, st = mu / a, ef = exp(-x[0] / a);
  double v = 2.0 * p->v0 * ef * sinh(st);
  double dSdpsi, dSdv;
  double S = rsfo_state_rate(p, i, x[0], v, &dSdpsi, &dSdv);
  double pre = 2.0 * p->v0 / (a * sig), cf = cosh(st) * ef;
  double dvdtau = pre * cf, dvdsig = -pre * cf * mu, dvdpsi = -v / a;
  double denom = 1.0 + E * dvdtau, Q = E * dvdpsi / denom;
  *fpsi = S;
  *ftau = -Q * S;
  if (J) {
    double vabs = fabs(v), sg = (v < 0.0) ? -1.0 : 1.0;
    /* total derivatives of S(psi, |v(psi, tau, sigma)|) */
    double Spsi = dSdpsi + dSdv * (-vabs / a);
    double Stau = dSdv * sg * dvdtau;
    double Ssig = dSdv * sg * dvdsig;
    /* derivatives of dv/dtau and of Q = E (dv/dpsi) / denom */
    double d2_psi = -dvdtau / a;
    double d2_tau = v / (a * a * sig * sig);
    double d2_sig = -dvdtau / sig - v * x[1] / (a * a * sig * sig * sig);
    double k = E / (denom * denom);
    double Qpsi = k * ((-dvdpsi / a) * denom - dvdpsi * E * d2_psi);
    double Qtau = k * ((-dvdtau / a) * denom - dvdpsi * E * d2_tau);
    double Qsig = k * ((-dvdsig / a) * denom - dvdpsi * E * d2_sig);
    memset(J, 0, 16 * sizeof(double));
    J[0] = Spsi; J[1] = Stau; J[2] = Ssig;
    J[4] = -(Q * Spsi + S * Qpsi);
    J[5] = -(Q * Stau + S * Qtau);
    J[6] = -(Q * Ssig + S * Qsig);
  }
}

void rsfo_imex_block(const rsfo_problem *p, int i, const double *x, double *J)
{
  double fp, ft;
  imex_local(p, i, x, &fp, &ft, J);
}

/* Stage equation of one cell, Y = Z + hg Fimpl(Y): sigma and slip are
   explicit (Y = Z); (psi, tau) solve R(u) = u - Z - hg Fimpl(u) = 0 by
   Newton's method with the exact 2x2 Jacobian I - hg dFimpl/du and
   backtracking on the weighted residual.  Converged when the weighted
   residual or the weighted update is at most IMEX_NEWTON_TOL (weights
   1/(atol + rtol |Z|), the step controller's).  The reference's PETSc SNES
   (newtonls, bt line search, rtol 1e-8 on the global residual 2-norm) is
   replaced by this per-cell test; both leave stage errors far below the
   step tolerance.  Returns the iteration count, or -1 on failure.  y holds
   the initial guess (psi, tau) on entry and the stage state on exit, fi
   the implicit stage derivative (Y - Z) / hg. */
#define IMEX_NEWTON_TOL 1e-6
#define IMEX_NEWTON_MAXIT 50
int rsfo_imex_newton(const rsfo_problem *p, int i, const double *Z, double hg,
                     double rtol, const double *atol, double *y, double *fi)
{
  double w0 = 1.0 / (atol[0] + rtol * fabs(Z[0])), w1 = 1.0 / (atol[1] + rtol * fabs(Z[1]));
  double u[4] = {y[0], y[1], Z[2], Z[3]}, J[16], fp, ft, r0, r1, rn;
  int it, conv = 0;
  imex_local(p, i, u, &fp, &ft, J);
  r0 = u[0] - Z[0] - hg * fp;
  r1 = u[1] - Z[1] - hg * ft;
  rn = fmax(fabs(r0) * w0, fabs(r1) * w1);
  for (it = 0; it < IMEX_NEWTON_MAXIT; it++) {
    double m00, m01, m10, m11, det, d0, d1, lam = 1.0, ut[4], rt0 = 0, rt1 = 0, rtn = 0;
    int k, ok = 0;
    if (!isfinite(rn)) return -1;
    if (rn <= IMEX_NEWTON_TOL) { conv = 1; break; }
    m00 = 1.0 - hg * J[0]; m01 = -hg * J[1];
    m10 = -hg * J[4]; m11 = 1.0 - hg * J[5];
    det = m00 * m11 - m01 * m10;
    if (!(det != 0.0) || !isfinite(det)) return -1;
    d0 = (-r0 * m11 + r1 * m01) / det;
    d1 = (-r1 * m00 + r0 * m10) / det;
    for (k = 0; k < 12; k++) {
      ut[0] = u[0] + lam * d0; ut[1] = u[1] + lam * d1; ut[2] = u[2]; ut[3] = u[3];
      imex_local(p, i, ut, &fp, &ft, J);
      rt0 = ut[0] - Z[0] - hg * fp;
      rt1 = ut[1] - Z[1] - hg * ft;
      rtn = fmax(fabs(rt0) * w0, fabs(rt1) * w1);
      if (isfinite(rtn) && rtn < (1.0 - 1e-4 * lam) * rn) { ok = 1; break; }
      lam *= 0.5;
    }
    if (!ok) return -1;
    u[0] = ut[0]; u[1] = ut[1];
    r0 = rt0; r1 = rt1; rn = rtn;
    if (fmax(fabs(lam * d0) * w0, fabs(lam * d1) * w1) <= IMEX_NEWTON_TOL) { conv = 1; it++; break; }
  }
  if (!conv) return -1;
  y[0] = u[0]; y[1] = u[1]; y[2] = Z[2]; y[3] = Z[3];
  fi[0] = (u[0] - Z[0]) / hg;
  fi[1] = (u[1] - Z[1]) / hg;
  fi[2] = 0.0;
  fi[3] = 0.0;
  return it;
}

static int is_normal(double v) { return isfinite(v) && (fpclassify(v) == FP_NORMAL); }

/* rsf_domain_check, src/rsf_engine.c:309-362 */
int rsfo_domain_ok(const rsfo_problem *p, const double *X)
{
  int i, dim = state_dim(p), nc = p->ncomp, k;
  for (i = 0; i < p->n; i++) {
    const double *x = X + (size_t)i * dim;
    double sig = x[nc + 1], tm;
    if (!is_normal(sig) || sig <= 0.0) return 0;
    for (k = 0; k < dim; k++)
      if (k != nc + 1 && !isfinite(x[k])) return 0;
    tm = (nc == 1) ? x[1] : sqrt(x[1] * x[1] + x[2] * x[2]);
    if (fabs(tm / (sig * p->a[i])) > ARG_MAX) return 0;
    if (fabs(x[0] / p->a[i]) > ARG_MAX) return 0;
    if (p->law != 2 && p->b[i] != 0.0 && fabs((p->f0 - x[0]) / p->b[i]) > ARG_MAX) return 0;
  }
  return 1;
}

/* ---------------- Runge-Kutta tableaus ---------------- */
typedef struct {
  int s, order, imex;
  double A[7][7], b[7], bh[7], c[7];
  double At[7][7];      /* IMEX: implicit tableau (A is the explicit one) */
  double d1[7], d2[7];  /* IMEX: dense output weights of theta and theta^2 */
} rk_tab;

static void tab_init(rk_tab *t, int method)
{
  memset(t, 0, sizeof(*t));
  if (method == 2) {
    /* PETSc 3.19.6 TSARKIMEXL2 (src/ts/impls/arkimex/arkimex.c, the
       reference's -imex default, src/rsf_imex.c:283): L-stable second
       order additive RK, explicit first stage, gamma = 1 - 1/sqrt 2;
       b = bt = last implicit row, embedded (first order) bembedt, and the
       quadratic dense output binterpt */
    const double s2 = sqrt(2.0), g = 1.0 - 1.0 / s2;
    t->s = 3; t->order = 2; t->imex = 1;
    t->c[1] = 2.0 - s2; t->c[2] = 1.0;
    t->A[1][0] = 2.0 - s2;
    t->A[2][0] = 0.5; t->A[2][1] = 0.5;
    t->At[1][0] = g; t->At[1][1] = g;
    t->At[2][0] = 1.0 / (2.0 * s2); t->At[2][1] = 1.0 / (2.0 * s2); t->At[2][2] = g;
    t->b[0] = t->At[2][0]; t->b[1] = t->At[2][1]; t->b[2] = t->At[2][2];
    t->bh[0] = (4.0 - s2) / 8.0; t->bh[1] = (4.0 - s2) / 8.0; t->bh[2] = 1.0 / (2.0 * s2);
    t->d1[0] = 1.0 / s2; t->d1[1] = 1.0 / s2; t->d1[2] = 1.0 - s2;
    t->d2[0] = -1.0 / (2.0 * s2); t->d2[1] = -1.0 / (2.0 * s2); t->d2[2] = 1.0 / s2;
  } else if (method == 1) { /* Bogacki-Shampine 3(2), FSAL */
    t->s = 4; t->order = 3;
    t->c[1] = 0.5; t->c[2] = 0.75; t->c[3] = 1.0;
    t->A[1][0] = 0.5;
    t->A[2][1] = 0.75;
    t->A[3][0] = 2.0 / 9.0; t->A[3][1] = 1.0 / 3.0; t->A[3][2] = 4.0 / 9.0;
    t->b[0] = 2.0 / 9.0; t->b[1] = 1.0 / 3.0; t->b[2] = 4.0 / 9.0; t->b[3] = 0.0;
    t->bh[0] = 7.0 / 24.0; t->bh[1] = 0.25; t->bh[2] = 1.0 / 3.0; t->bh[3] = 0.125;
  } else { /* Dormand-Prince 5(4), FSAL (Dormand & Prince 1980) */
    t->s = 7; t->order = 5;
    t->c[1] = 0.2; t->c[2] = 0.3; t->c[3] = 0.8; t->c[4] = 8.0 / 9.0; t->c[5] = 1.0; t->c[6] = 1.0;
    t->A[1][0] = 0.2;
    t->A[2][0] = 3.0 / 40.0; t->A[2][1] = 9.0 / 40.0;
    t->A[3][0] = 44.0 / 45.0; t->A[3][1] = -56.0 / 15.0; t->A[3][2] = 32.0 / 9.0;
    t->A[4][0] = 19372.0 / 6561.0; t->A[4][1] = -25360.0 / 2187.0; t->A[4][2] = 64448.0 / 6561.0; t->A[4][3] = -212.0 / 729.0;
    t->A[5][0] = 9017.0 / 3168.0; t->A[5][1] = -355.0 / 33.0; t->A[5][2] = 46732.0 / 5247.0; t->A[5][3] = 49.0 / 176.0; t->A[5][4] = -5103.0 / 18656.0;
    t->A[6][0] = 35.0 / 384.0; t->A[6][2] = 500.0 / 1113.0; t->A[6][3] = 125.0 / 192.0; t->A[6][4] = -2187.0 / 6784.0; t->A[6][5] = 11.0 / 84.0;
    memcpy(t->b, t->A[6], sizeof(t->b));
    t->bh[0] = 5179.0 / 57600.0; t->bh[2] = 7571.0 / 16695.0; t->bh[3] = 393.0 / 640.0;
    t->bh[4] = -92097.0 / 339200.0; t->bh[5] = 187.0 / 2100.0; t->bh[6] = 1.0 / 40.0;
  }
}

/* dense output on [t, t+h] at fraction th (Hairer's DOPRI5 continuous
   extension; Hermite cubic for 3bs) */
static void dense_eval(const rk_tab *t, int m, double h, const double *X0,
                       const double *X1, double *const *K, double th, double *out)
{
  int i;
  if (t->imex) {
    /* K[j] = G_j, K[3 + j] = Fimpl_j */
    for (i = 0; i < m; i++) {
      double acc = 0.0;
      int j;
      for (j = 0; j < 3; j++)
        acc += (t->d1[j] * th + t->d2[j] * th * th) * (K[j][i] + K[3 + j][i]);
      out[i] = X0[i] + h * acc;
    }
    return;
  }
  if (t->s == 7) {
    static const double d1 = -12715105075.0 / 11282082432.0, d3 = 87487479700.0 / 32700410799.0,
                        d4 = -10690763975.0 / 1880347072.0, d5 = 701980252875.0 / 199316789632.0,
                        d6 = -1453857185.0 / 822651844.0, d7 = 69997945.0 / 29380423.0;
    for (i = 0; i < m; i++) {
      double r2 = X1[i] - X0[i];
      double r3 = h * K[0][i] - r2;
      double r4 = r2 - h * K[6][i] - r3;
      double r5 = h * (d1 * K[0][i] + d3 * K[2][i] + d4 * K[3][i] + d5 * K[4][i] + d6 * K[5][i] + d7 * K[6][i]);
      out[i] = X0[i] + th * (r2 + (1.0 - th) * (r3 + th * (r4 + (1.0 - th) * r5)));
    }
  } else {
    for (i = 0; i < m; i++) {
      double t2 = th * th, t3 = t2 * th;
      out[i] = (2 * t3 - 3 * t2 + 1) * X0[i] + (t3 - 2 * t2 + th) * h * K[0][i] +
               (-2 * t3 + 3 * t2) * X1[i] + (t3 - t2) * h * K[3][i];
    }
  }
}

typedef struct {
  double vmax, vsum, slip, mu, smax, smin;
} field_stats;

static void stats(const rsfo_problem *p, const double *X, field_stats *s)
{
  int i, dim = state_dim(p), nc = p->ncomp;
  double v[2];
  s->vmax = 0.0; s->vsum = 0.0; s->slip = 0.0; s->mu = 0.0; s->smax = -1e30; s->smin = 1e30;
  for (i = 0; i < p->n; i++) {
    const double *x = X + (size_t)i * dim;
    double va;
    patch_vel(p, i, x, v);
    va = (nc == 1) ? fabs(v[0]) : sqrt(v[0] * v[0] + v[1] * v[1]);
    if (va > s->vmax) s->vmax = va;
    s->vsum += va;
    s->slip += x[nc + 2];
    s->mu += x[1] / x[nc + 1];
    if (x[nc + 1] > s->smax) s->smax = x[nc + 1];
    if (x[nc + 1] < s->smin) s->smin = x[nc + 1];
  }
}

static double evfun(double vmax, double thr)
{
  return log10((vmax > 1e-300) ? vmax : 1e-300) - log10(thr);
}

int rsfo_rk_attempt(const rsfo_problem *p, int method, double rtol,
                    const double *atol, double h, const double *X,
                    const double *F0, double *Xnew, double *Fnew, double *errnorm)
{
  rk_tab tb;
  int dim = state_dim(p), m = dim * p->n, s, j, i, ok = 1;
  double *K[7] = {0}, *Y = (double *)malloc(sizeof(double) * m), en = 0.0;
  tab_init(&tb, method);
  for (s = 0; s < tb.s; s++) K[s] = (double *)malloc(sizeof(double) * m);
  memcpy(K[0], F0, sizeof(double) * m);
  for (s = 1; s < tb.s; s++) {
    for (i = 0; i < m; i++) {
      double acc = 0.0;
      for (j = 0; j < s; j++) acc += tb.A[s][j] * K[j][i];
      Y[i] = X[i] + h * acc;
    }
    if (!rsfo_domain_ok(p, Y)) ok = 0;
    rsfo_rhs(p, Y, K[s], NULL, NULL);
  }
  for (i = 0; i < m; i++) {
    double a5 = 0.0, a4 = 0.0, tol;
    for (j = 0; j < tb.s; j++) { a5 += tb.b[j] * K[j][i]; a4 += tb.bh[j] * K[j][i]; }
    Xnew[i] = X[i] + h * a5;
    a4 = X[i] + h * a4;
    tol = atol[i % dim] + rtol * fmax(fabs(Xnew[i]), fabs(a4));
    en = fmax(en, fabs(Xnew[i] - a4) / tol);
  }
  memcpy(Fnew, K[tb.s - 1], sizeof(double) * m);
  *errnorm = en;
  for (s = 0; s < tb.s; s++) free(K[s]);
  free(Y);
  return ok;
}

/* the accepted-step loop of TSSolve with the rsf monitor and event hooks
   (src/rsf_solve.c:440-643, src/rsf_output.c:99-198, 718-864, 1000-1150) */
int rsfo_solve(const rsfo_problem *p, const rsfo_options *o, double *X,
               rsfo_result *r)
{
  rk_tab tb;
  int dim = state_dim(p), m = dim * p->n, n = p->n, nc = p->ncomp, s, j, i;
  double *K[7], *Y, *Xn, *Xmon, *Xe, *tau0, *slip0, *work;
  unsigned char *rupt;
  double t = o->t0, h = o->dt_init, hmax, t_old_mon, thr;
  int slipping, storm = 0, nrup, ev_open = 0;
  double fprev, peakv = 0.0, t_onset = 0.0;
  field_stats st;
  tab_init(&tb, o->method);
  if (tb.imex && nc != 1) {       /* the reference's IMEX is scalar-slip only */
    r->steps = r->rejects = r->domain_rejects = r->rhs_evals = 0;
    r->nevent = r->nmon = 0;
    r->status = 4;
    r->t_final = t;
    r->dt_final = h;
    return 4;
  }
  for (s = 0; s < 7; s++) K[s] = (double *)malloc(sizeof(double) * m);
  Y = (double *)malloc(sizeof(double) * m);
  Xn = (double *)malloc(sizeof(double) * m);
  Xmon = (double *)malloc(sizeof(double) * m);
  Xe = (double *)malloc(sizeof(double) * m);
  work = (double *)malloc(sizeof(double) * m);
  tau0 = (double *)calloc(n, sizeof(double));
  slip0 = (double *)calloc(n, sizeof(double));
  rupt = (unsigned char *)calloc(n, 1);
  r->steps = r->rejects = r->domain_rejects = r->rhs_evals = 0;
  r->status = 0;
  r->nevent = 0;
  r->nmon = 0;
  r->ncat = 0;
  hmax = (o->dt_max < o->dt_monitor) ? o->dt_max : o->dt_monitor;
  if (tb.imex) rsfo_imex_parts(p, X, K[0], K[3]);
  else rsfo_rhs(p, X, K[0], NULL, NULL);
  r->rhs_evals++;
  stats(p, X, &st);
  slipping = st.vmax > o->vel_event;
  thr = o->vel_event * (slipping ? o->vel_event_hyst : 1.0);
  fprev = evfun(st.vmax, thr);
  memcpy(Xmon, X, sizeof(double) * m);
  t_old_mon = o->t0 - 2.0 * o->dt_monitor;
  for (;;) {
    int accept_flag = 1, accepted = 0;
    double enorm = 0.0;
    /* ---- monitor at the current accepted state (TSMonitor) ---- */
    {
      double xn2 = 0.0, dx2 = 0.0;
      int bail = 0;
      for (i = 0; i < m; i++) {
        xn2 += X[i] * X[i];
        dx2 += (Xmon[i] - X[i]) * (Xmon[i] - X[i]);
      }
      xn2 = sqrt(xn2);
      dx2 = sqrt(dx2);
      if (xn2 > 1e-15 && dx2 / xn2 > o->rdx_monitor) bail = 1;
      if (o->adx_monitor > 0 && dx2 > o->adx_monitor) bail = 1;
      if (fabs(t - t_old_mon) > o->dt_monitor) bail = 1;
      if (bail) {
        stats(p, X, &st);
        if (r->mon && r->nmon < o->max_monitor) {
          double *row = r->mon + 10 * r->nmon;
          row[0] = (double)r->steps; row[1] = t; row[2] = t / SEC_PER_YEAR; row[3] = h;
          row[4] = log10(st.vmax); row[5] = st.slip / n; row[6] = st.mu / n;
          row[7] = st.smax; row[8] = st.smin;
          row[9] = log10((st.vsum / n > 1e-300) ? st.vsum / n : 1e-300);
          r->nmon++;
        }
        t_old_mon = t;
        memcpy(Xmon, X, sizeof(double) * m);
      }
    }
    if (!(t < o->t_end)) break;
    if (r->steps >= o->max_steps) { r->status = 2; break; }
    /* ---- one TSStep: attempts until accepted ---- */
    while (!accepted) {
      int stage_ok = 1;
      /* stage 0 is the accepted state; its check resets the storm counter */
      if (rsfo_domain_ok(p, X)) storm = 0;
      else { r->status = 3; goto done; }
      if (tb.imex) {
        /* ARKIMEX stages 2 and 3: Z = X + h sum_j (At G-free implicit + A explicit),
           per-cell Newton for Y = Z + h gamma Fimpl(Y) from the previous
           stage's state (TSStep_ARKIMEX), then G at Y (one matvec) */
        double *Yp = Xe;        /* previous stage state (initial guess) */
        memcpy(Yp, X, sizeof(double) * m);
        for (s = 1; s < tb.s && stage_ok; s++) {
          const double hg = h * tb.At[s][s];
          for (i = 0; i < n && stage_ok; i++) {
            double Z[4], yv[4];
            int f;
            for (f = 0; f < 4; f++) {
              size_t q = 4 * (size_t)i + f;
              double acc = 0.0;
              for (j = 0; j < s; j++) acc += tb.At[s][j] * K[3 + j][q] + tb.A[s][j] * K[j][q];
              Z[f] = X[q] + h * acc;
            }
            yv[0] = Yp[4 * (size_t)i]; yv[1] = Yp[4 * (size_t)i + 1];
            if (rsfo_imex_newton(p, i, Z, hg, o->rtol, o->atol, yv, K[3 + s] + 4 * (size_t)i) < 0)
              stage_ok = 0;
            for (f = 0; f < 4; f++) Y[4 * (size_t)i + f] = yv[f];
          }
          if (stage_ok && !rsfo_domain_ok(p, Y)) stage_ok = 0;
          if (!stage_ok) break;
          rsfo_imex_parts(p, Y, K[s], work);
          r->rhs_evals++;
          memcpy(Yp, Y, sizeof(double) * m);
        }
        if (stage_ok)
          for (i = 0; i < m; i++) {
            double acc = 0.0;
            for (j = 0; j < tb.s; j++) acc += tb.b[j] * (K[j][i] + K[3 + j][i]);
            Xn[i] = X[i] + h * acc;
          }
      }
      for (s = 1; s < tb.s && stage_ok && !tb.imex; s++) {
        for (i = 0; i < m; i++) {
          double acc = 0.0;
          for (j = 0; j < s; j++) acc += tb.A[s][j] * K[j][i];
          Y[i] = X[i] + h * acc;
        }
        if (!rsfo_domain_ok(p, Y)) {
          stage_ok = 0;
          break;
        }
        rsfo_rhs(p, Y, K[s], NULL, NULL);
        r->rhs_evals++;
        if (s == tb.s - 1) memcpy(Xn, Y, sizeof(double) * m);
      }
      if (!stage_ok) {
        r->domain_rejects++;
        r->rejects++;
        storm++;
        if (o->max_domain_reject > 0 && storm >= o->max_domain_reject) {
          r->status = 1;
          goto done;
        }
        h *= 0.25;                    /* TSAdapt scale_solve_failed */
        accept_flag = 0;
        continue;
      }
      /* embedded solution and weighted infinity norm (TSErrorWeightedNorm) */
      enorm = 0.0;
      for (i = 0; i < m; i++) {
        double a4 = 0.0, tol;
        for (j = 0; j < tb.s; j++) a4 += tb.bh[j] * (tb.imex ? K[j][i] + K[3 + j][i] : K[j][i]);
        a4 = X[i] + h * a4;
        tol = o->atol[i % dim] + o->rtol * fmax(fabs(Xn[i]), fabs(a4));
        if (tol > 0.0) enorm = fmax(enorm, fabs(Xn[i] - a4) / tol);
      }
      if (!isfinite(enorm)) enorm = 1e300;
      {
        double safety = 0.9, hfac, hnew;
        int acc = 1;
        if (enorm > 1.0) {
          if (!accept_flag) safety *= 0.5;
          acc = 0;
        }
        hfac = (enorm > 0.0) ? safety * pow(enorm, -1.0 / tb.order) : 2.0;
        if (hfac < 0.5) hfac = 0.5;
        if (hfac > 2.0) hfac = 2.0;
        hnew = h * hfac;
        if (hnew > hmax) hnew = hmax;
        if (!acc) {
          h = hnew;
          accept_flag = 0;
          r->rejects++;
          continue;
        }
        /* accepted */
        accepted = 1;
        /* events on the dense output of [t, t + h] */
        {
          double told = t, hstep = h;
          double th_lo = 0.0;
          field_stats se;
          stats(p, Xn, &se);
          for (;;) {
            double fnew = evfun(se.vmax, thr);
            double a_th, b_th, fa, fb;
            int side = 0, it;
            if (!((fprev < 0.0 && fnew > 0.0) || (fprev > 0.0 && fnew < 0.0) || fnew == 0.0)) {
              fprev = fnew;
              break;
            }
            /* Illinois regula falsi on [th_lo, 1] */
            a_th = th_lo; fa = fprev; b_th = 1.0; fb = fnew;
            for (it = 0; it < 200; it++) {
              double c_th = (fb - fa != 0.0) ? b_th - fb * (b_th - a_th) / (fb - fa) : 0.5 * (a_th + b_th);
              double fc;
              if (!(c_th > a_th && c_th < b_th)) c_th = 0.5 * (a_th + b_th);
              dense_eval(&tb, m, hstep, X, Xn, K, c_th, Xe);
              stats(p, Xe, &st);
              fc = evfun(st.vmax, thr);
              if (fabs(fc) < 1e-10 || (b_th - a_th) * hstep < 1e-9) { a_th = b_th = c_th; break; }
              if ((fc > 0.0) == (fb > 0.0)) {
                b_th = c_th; fb = fc;
                if (side == -1) fa *= 0.5;
                side = -1;
              } else {
                a_th = c_th; fa = fc;
                if (side == 1) fb *= 0.5;
                side = 1;
              }
            }
            dense_eval(&tb, m, hstep, X, Xn, K, b_th, Xe);
            stats(p, Xe, &st);
            /* rsf_post_event: toggle, log, catalog snapshots */
            slipping = !slipping;
            if (slipping) {
              peakv = st.vmax;
              for (i = 0; i < n; i++) {
                tau0[i] = Xe[(size_t)i * dim + 1];
                slip0[i] = Xe[(size_t)i * dim + nc + 2];
                rupt[i] = 0;
              }
            }
            if (r->ev && r->nevent < o->max_events) {
              double *row = r->ev + 8 * r->nevent;
              double drop = 0.0;
              nrup = 0;
              if (!slipping) {
                for (i = 0; i < n; i++)
                  if (rupt[i]) {
                    drop += fabs(tau0[i]) - fabs(Xe[(size_t)i * dim + 1]);
                    nrup++;
                  }
              }
              row[0] = told + b_th * hstep;
              row[1] = slipping ? 1.0 : -1.0;
              row[2] = log10((st.vmax > 1e-300) ? st.vmax : 1e-300);
              row[3] = st.slip / n;
              row[4] = st.mu / n;
              row[5] = slipping ? st.vmax : fmax(peakv, st.vmax);
              row[6] = nrup ? drop / nrup : 0.0;
              row[7] = nrup;
              r->nevent++;
            }
            /* arrest of an open event: its catalog row when a cell
               ruptured (rsf_finalize_event, src/rsf_output.c:1090-1178):
               |slip change| and change of |tau| since the onset snapshot,
               area 4 l w, M0 = mu sum area |slip change|, Mw from
               mwfromm0 (src/geometry/llgeo.c:757-760) */
            if (!slipping) {
              double sds = 0.0, mds = 0.0, sdd = 0.0, mdd = 0.0, sar = 0.0, smo = 0.0;
              int cnt = 0;
              for (i = 0; i < n; i++)
                if (rupt[i]) {
                  const double dd = fabs(tau0[i]) - fabs(Xe[(size_t)i * dim + 1]);
                  const double ds = fabs(Xe[(size_t)i * dim + nc + 2] - slip0[i]);
                  const double ar = r->cat_area ? r->cat_area[i] : 0.0;
                  sds += ds; sdd += dd; sar += ar; smo += ar * ds;
                  if (ds > mds) mds = ds;
                  if (dd > mdd) mdd = dd;
                  cnt++;
                }
              if (ev_open && cnt > 0 && r->cat && r->ncat < o->max_events) {
                double *row = r->cat + 11 * r->ncat;
                const double m0 = r->cat_mu * smo;
                row[0] = t_onset;
                row[1] = told + b_th * hstep;
                row[2] = cnt;
                row[3] = sar;
                row[4] = sds / cnt;
                row[5] = mds;
                row[6] = sdd / cnt;
                row[7] = mdd;
                row[8] = fmax(peakv, st.vmax);
                row[9] = m0;
                row[10] = (m0 > 0.0) ? (2.0 / 3.0) * (log10(m0) - 9.05) : -99.0;
                r->ncat++;
              }
            } else {
              t_onset = told + b_th * hstep;
            }
            ev_open = slipping;
            thr = o->vel_event * (slipping ? o->vel_event_hyst : 1.0);
            fprev = evfun(st.vmax, thr);
            th_lo = b_th;
            if (th_lo >= 1.0) {
              fprev = evfun(se.vmax, thr);
              break;
            }
          }
        }
        if (r->hlog && r->steps < r->hlog_cap) r->hlog[r->steps] = h;
        t += h;
        h = hnew;
        memcpy(X, Xn, sizeof(double) * m);
        if (tb.imex) {                /* stage 1 of the next step: G and Fimpl at X */
          rsfo_imex_parts(p, X, K[0], K[3]);
          r->rhs_evals++;
        } else {
          memcpy(K[0], K[tb.s - 1], sizeof(double) * m);   /* FSAL */
        }
      }
    }
    r->steps++;
    /* per-step catalog sampling while slipping (rsf_output.c:737-775) */
    if (slipping) {
      for (i = 0; i < n; i++) {
        double v[2], va;
        patch_vel(p, i, X + (size_t)i * dim, v);
        va = (nc == 1) ? fabs(v[0]) : sqrt(v[0] * v[0] + v[1] * v[1]);
        if (va > peakv) peakv = va;
        if (va >= o->rupture_vth) rupt[i] = 1;
      }
    }
  }
done:
  r->t_final = t;
  r->dt_final = h;
  for (s = 0; s < 7; s++) free(K[s]);
  free(Y); free(Xn); free(Xmon); free(Xe); free(work);
  free(tau0); free(slip0); free(rupt);
  return r->status;
}
