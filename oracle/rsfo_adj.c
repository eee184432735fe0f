/*
  Oracle restatement of the discrete adjoint of the explicit Runge-Kutta
  cycle integration (the FP64 reference adjoint).

  TEST INFRASTRUCTURE ONLY (parity oracle).

  The reference has no adjoint.  The differentiated map is a replay of the
  rsfo_rk_attempt step (rk5dp or rk3bs stages over rsfo_rhs, the
  restatement of rsf_ODE_RHSFunction, src/rsf_engine.c:186-302) on a fixed
  list of step sizes, and the objective
    J = 1/2 sum_n weight[n] sum_r (ln|v_{recv[r]}(X_n)| - target[n][r])^2.
  Local (per patch) Jacobians of the right-hand side, of the slip rate and
  of ln|v| are taken by complex-step differentiation (Squire & Trapp 1998,
  SIAM Review 40:110; Martins, Sturdza & Alonso 2003, ACM TOMS 29:245):
  f'(x) = Im f(x + i e) / e with e = 1e-30, exact to rounding for the
  analytic law, on a complex restatement of the same formulas.  The
  backward sweep stores every stage state (small problems only) and applies
  K^T by plain loops.
*/
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "rsfo.h"

typedef double complex cplx;

#define CS_STEP 1e-30

static cplx c_absr(cplx z) { return (creal(z) < 0.0) ? -z : z; }

/* rsf_state_rate (src/rsf_engine.c:44-132) in complex arithmetic */
static cplx state_rate_c(const rsfo_problem *p, cplx b, cplx D, cplx psi, cplx vabs)
{
  cplx e, om, gate, lom, pss;
  if (creal(b) == 0.0) return 0.0;
  switch (p->law) {
  case 2:
    if (creal(vabs) < p->vmin_state) vabs = p->vmin_state;
    pss = p->f0 - b * clog(vabs / p->v0);
    return -(vabs / D) * (psi - pss);
  case 3:
    e = cexp((psi - p->f0) / b);
    return (b / (2.0 * D)) * (p->v0 / e - (vabs * vabs / p->v0) * e);
  case 4:
  case 5:
    if (creal(vabs) < p->vmin_state) vabs = p->vmin_state;
    e = cexp((psi - p->f0) / b);
    om = (vabs / p->v0) * e;
    gate = (p->law == 4) ? cexp(-om / p->sato_beta) : cexp(-vabs / p->kt_vc);
    lom = clog(vabs / p->v0) + (psi - p->f0) / b;
    return (b / D) * ((p->v0 / e) * gate - vabs * lom);
  default:
    e = cexp((p->f0 - psi) / b);
    return (b / D) * (p->v0 * e - vabs);
  }
}

/* slip rate components of one patch (vel_from_rsf, src/rsf_engine.c:20-30) */
static void vel_c(const rsfo_problem *p, cplx a, const cplx *x, cplx *v)
{
  if (p->ncomp == 1) {
    v[0] = 2.0 * p->v0 * cexp(-x[0] / a) * csinh((x[1] / x[2]) / a);
  } else {
    cplx tm = csqrt(x[1] * x[1] + x[2] * x[2]);
    cplx vm = 2.0 * p->v0 * cexp(-x[0] / a) * csinh(tm / (x[3] * a));
    if (creal(tm) > 0.0) {
      v[0] = vm * x[1] / tm;
      v[1] = vm * x[2] / tm;
    } else {
      v[0] = v[1] = 0.0;
    }
  }
}

/* the right-hand side of patch i from its state x and stressing rates td
   (rsfo_rhs per patch, without normal-stress rows) */
static void rhs_c(const rsfo_problem *p, int i, cplx a, cplx b, cplx dc, const cplx *x,
                  const cplx *td, cplx *f)
{
  int nc = p->ncomp;
  double eta = p->eta, sdot;
  cplx v[2], vmag, sig = x[nc + 1];
  vel_c(p, a, x, v);
  vmag = (nc == 1) ? v[0] : csqrt(v[0] * v[0] + v[1] * v[1]);
  f[0] = state_rate_c(p, b, dc, x[0], c_absr(vmag));
  sdot = p->calc_sigma_dot ? (p->sinc_n ? p->sinc_n[i] : 0.0) : 0.0;
  if (p->limit_sigma && ((creal(sig) < p->min_sigma) || (creal(sig) > p->max_sigma))) sdot = 0.0;
  f[nc + 1] = sdot;
  if (nc == 1) {
    cplx mu = x[1] / x[2], st = mu / a, ef = cexp(-x[0] / a);
    cplx pre = 2.0 * p->v0 / (a * x[2]), cf = ccosh(st) * ef;
    cplx dvdtau = pre * cf, dvdsig = -pre * cf * mu, dvdpsi = -v[0] / a;
    f[1] = (td[0] + p->sinc[i] - eta * (dvdsig * f[2] + dvdpsi * f[0])) / (1.0 + eta * dvdtau);
    f[3] = v[0];
  } else {
    cplx ts = x[1], tdp = x[2], tm = csqrt(ts * ts + tdp * tdp);
    cplx arg = tm / (sig * a), ef = cexp(-x[0] / a);
    cplx gpar = 2.0 * p->v0 / (a * sig) * ccosh(arg) * ef;
    cplx gperp = (creal(arg) > 1e-8) ? vmag / tm : 2.0 * p->v0 / (a * sig) * ef;
    cplx th0 = (creal(tm) > 0.0) ? ts / tm : 1.0, th1 = (creal(tm) > 0.0) ? tdp / tm : 0.0;
    cplx dvmdsig = -gpar * tm / sig, r0, r1, rp;
    r0 = td[0] + p->sinc[2 * i] - eta * (dvmdsig * th0 * f[3] - v[0] / a * f[0]);
    r1 = td[1] + p->sinc[2 * i + 1] - eta * (dvmdsig * th1 * f[3] - v[1] / a * f[0]);
    rp = th0 * r0 + th1 * r1;
    f[1] = th0 * rp / (1.0 + eta * gpar) + (r0 - th0 * rp) / (1.0 + eta * gperp);
    f[2] = th1 * rp / (1.0 + eta * gpar) + (r1 - th1 * rp) / (1.0 + eta * gperp);
    f[4] = v[0];
    f[5] = v[1];
  }
}

static int dim_of(const rsfo_problem *p) { return p->ncomp == 2 ? 6 : 4; }

void rsfo_patch_rhs_c(const rsfo_problem *p, int i, const double *x, const double *td,
                      double *f, double *v)
{
  int dim = dim_of(p), nc = p->ncomp, k;
  cplx xc[6], tc[2], fc[6], vc[2];
  for (k = 0; k < dim; k++) xc[k] = x[k];
  for (k = 0; k < nc; k++) tc[k] = td[k];
  rhs_c(p, i, p->a[i], p->b[i], p->dc[i], xc, tc, fc);
  vel_c(p, p->a[i], xc, vc);
  for (k = 0; k < dim; k++) f[k] = creal(fc[k]);
  for (k = 0; k < nc; k++) v[k] = creal(vc[k]);
}

/* Jacobians of patch i by complex step: Jx[dim][dim] = df/dx,
   Jtd[dim][nc] = df/dtd, Jp[dim][3] = df/d(a, b, dc), Vx[nc][dim] = dv/dx,
   Vp[nc] = dv/da, Lx[dim] = d ln|v| / dx, *La = d ln|v| / da (row-major;
   any output may be NULL) */
void rsfo_patch_jacobian(const rsfo_problem *p, int i, const double *x, const double *td,
                         double *Jx, double *Jtd, double *Jp, double *Vx, double *Vp,
                         double *Lx, double *La)
{
  int dim = dim_of(p), nc = p->ncomp, q, k, f;
  const double h = CS_STEP;
  for (q = 0; q < dim + nc + 3; q++) {
    cplx xc[6], tc[2], fc[6], vc[2], vm, ca = p->a[i], cb = p->b[i], cd = p->dc[i];
    for (k = 0; k < dim; k++) xc[k] = x[k];
    for (k = 0; k < nc; k++) tc[k] = td[k];
    if (q < dim) xc[q] += I * h;
    else if (q < dim + nc) tc[q - dim] += I * h;
    else if (q == dim + nc) ca += I * h;
    else if (q == dim + nc + 1) cb += I * h;
    else cd += I * h;
    rhs_c(p, i, ca, cb, cd, xc, tc, fc);
    vel_c(p, ca, xc, vc);
    vm = (nc == 1) ? c_absr(vc[0]) : csqrt(vc[0] * vc[0] + vc[1] * vc[1]);
    for (f = 0; f < dim; f++) {
      double d = cimag(fc[f]) / h;
      if (q < dim) { if (Jx) Jx[f * dim + q] = d; }
      else if (q < dim + nc) { if (Jtd) Jtd[f * nc + (q - dim)] = d; }
      else if (Jp) Jp[f * 3 + (q - dim - nc)] = d;
    }
    for (k = 0; k < nc; k++) {
      double d = cimag(vc[k]) / h;
      if (q < dim) { if (Vx) Vx[k * dim + q] = d; }
      else if (q == dim + nc) { if (Vp) Vp[k] = d; }
    }
    if (q < dim) { if (Lx) Lx[q] = cimag(clog(vm)) / h; }
    else if (q == dim + nc) { if (La) *La = cimag(clog(vm)) / h; }
  }
}

static double ln_speed(const rsfo_problem *p, int i, const double *x)
{
  double f[6], v[2];
  double td[2] = {0.0, 0.0};
  rsfo_patch_rhs_c(p, i, x, td, f, v);
  return (p->ncomp == 1) ? log(fabs(v[0])) : log(sqrt(v[0] * v[0] + v[1] * v[1]));
}

/* the rk5dp / rk3bs tableau (the rows of rsfo_rk_attempt's stages) */
static int tableau(int method, double A[7][7])
{
  memset(A, 0, sizeof(double) * 49);
  if (method == 1) {
    A[1][0] = 0.5;
    A[2][1] = 0.75;
    A[3][0] = 2.0 / 9.0; A[3][1] = 1.0 / 3.0; A[3][2] = 4.0 / 9.0;
    return 4;
  }
  A[1][0] = 0.2;
  A[2][0] = 3.0 / 40.0; A[2][1] = 9.0 / 40.0;
  A[3][0] = 44.0 / 45.0; A[3][1] = -56.0 / 15.0; A[3][2] = 32.0 / 9.0;
  A[4][0] = 19372.0 / 6561.0; A[4][1] = -25360.0 / 2187.0; A[4][2] = 64448.0 / 6561.0; A[4][3] = -212.0 / 729.0;
  A[5][0] = 9017.0 / 3168.0; A[5][1] = -355.0 / 33.0; A[5][2] = 46732.0 / 5247.0; A[5][3] = 49.0 / 176.0; A[5][4] = -5103.0 / 18656.0;
  A[6][0] = 35.0 / 384.0; A[6][2] = 500.0 / 1113.0; A[6][3] = 125.0 / 192.0; A[6][4] = -2187.0 / 6784.0; A[6][5] = 11.0 / 84.0;
  return 7;
}

/* Replay: X_{m+1} = X_m + h sum_j A[S-1][j] K_j with the stages of
   rsfo_rk_attempt.  Optional: trace [(N+1) x nrecv] ln|v|, the stage states
   recY [N][S-1][m] and stressing rates recT [N][S-1][nc n]. */
static void replay_core(const rsfo_problem *p, int method, const double *X0,
                        const double *steps, long N, const int *recv, int nrecv,
                        double *trace, double *recY, double *recT, double *Xend)
{
  double A[7][7];
  int S = tableau(method, A), dim = dim_of(p), n = p->n, nc = p->ncomp;
  long m = (long)dim * n, k, step;
  int s, j, r;
  double *X = (double *)malloc(sizeof(double) * m), *Y = (double *)malloc(sizeof(double) * m);
  double *K[7], *td = (double *)malloc(sizeof(double) * nc * n);
  for (s = 0; s < S; s++) K[s] = (double *)malloc(sizeof(double) * m);
  memcpy(X, X0, sizeof(double) * m);
  rsfo_rhs(p, X, K[0], NULL, td);
  for (step = 0; step <= N; step++) {
    if (trace)
      for (r = 0; r < nrecv; r++)
        trace[step * nrecv + r] = ln_speed(p, recv[r], X + (size_t)recv[r] * dim);
    if (step == N) break;
    if (recY) {
      memcpy(recY + ((size_t)step * (S - 1)) * m, X, sizeof(double) * m);
      memcpy(recT + ((size_t)step * (S - 1)) * nc * n, td, sizeof(double) * nc * n);
    }
    const double h = steps[step];
    for (s = 1; s < S; s++) {
      for (k = 0; k < m; k++) {
        double acc = 0.0;
        for (j = 0; j < s; j++) acc += A[s][j] * K[j][k];
        Y[k] = X[k] + h * acc;
      }
      rsfo_rhs(p, Y, K[s], NULL, td);
      if (recY && s < S - 1) {
        memcpy(recY + ((size_t)step * (S - 1) + s) * m, Y, sizeof(double) * m);
        memcpy(recT + ((size_t)step * (S - 1) + s) * nc * n, td, sizeof(double) * nc * n);
      }
    }
    memcpy(X, Y, sizeof(double) * m);          /* Y_{S-1} = X_{m+1} */
    memcpy(K[0], K[S - 1], sizeof(double) * m); /* FSAL */
  }
  if (Xend) memcpy(Xend, X, sizeof(double) * m);
  for (s = 0; s < S; s++) free(K[s]);
  free(X); free(Y); free(td);
}

void rsfo_replay(const rsfo_problem *p, int method, const double *X0, const double *steps,
                 long nsteps, const int *recv, int nrecv, double *trace, double *Xend)
{
  replay_core(p, method, X0, steps, nsteps, recv, nrecv, trace, NULL, NULL, Xend);
}

double rsfo_adjoint(const rsfo_problem *p, int method, const double *X0, const double *steps,
                    long N, const int *recv, int nrecv, const double *target,
                    const double *weight, double *ga, double *gb, double *gdc)
{
  double A[7][7];
  int S = tableau(method, A), dim = dim_of(p), n = p->n, nc = p->ncomp, s, i, j, f, c, r;
  long m = (long)dim * n, k, step;
  int nd = nc * n;
  double *recY = (double *)malloc(sizeof(double) * (size_t)N * (S - 1) * m);
  double *recT = (double *)malloc(sizeof(double) * (size_t)N * (S - 1) * nd);
  double *trace = (double *)malloc(sizeof(double) * (size_t)(N + 1) * nrecv);
  double *Xe = (double *)malloc(sizeof(double) * m);
  double *lam = (double *)calloc(m, sizeof(double));
  double *M[7], *w = (double *)malloc(sizeof(double) * nd), *z = (double *)malloc(sizeof(double) * nd);
  double *nu = (double *)malloc(sizeof(double) * m);
  int *recv_of = (int *)malloc(sizeof(int) * n);
  double J = 0.0;
  if (p->Kn && p->calc_sigma_dot) return NAN;   /* normal-stress coupling: not restated */
  for (s = 0; s < 7; s++) M[s] = (double *)malloc(sizeof(double) * m);
  for (i = 0; i < n; i++) recv_of[i] = -1;
  for (r = 0; r < nrecv; r++) recv_of[recv[r]] = r;
  replay_core(p, method, X0, steps, N, recv, nrecv, trace, recY, recT, Xe);
  for (step = 0; step <= N; step++)
    for (r = 0; r < nrecv; r++) {
      double d = trace[step * nrecv + r] - target[step * nrecv + r];
      J += 0.5 * weight[step] * d * d;
    }
  for (i = 0; i < n; i++) ga[i] = gb[i] = gdc[i] = 0.0;
  /* misfit at state N */
  for (i = 0; i < n; i++) {
    double Lx[6], La, td0[2] = {0.0, 0.0};
    if (recv_of[i] < 0 || weight[N] == 0.0) continue;
    rsfo_patch_jacobian(p, i, Xe + (size_t)i * dim, td0, NULL, NULL, NULL, NULL, NULL, Lx, &La);
    {
      double res = trace[N * nrecv + recv_of[i]] - target[N * nrecv + recv_of[i]];
      for (f = 0; f < dim; f++) lam[(size_t)i * dim + f] += weight[N] * res * Lx[f];
      ga[i] += weight[N] * res * La;
    }
  }
  for (step = N - 1; step >= 0; step--) {
    const double h = steps[step];
    for (j = 0; j < S - 1; j++)
      for (k = 0; k < m; k++) M[j][k] = h * A[S - 1][j] * lam[k];
    for (s = S - 2; s >= 0; s--) {
      const double *Ys = recY + ((size_t)step * (S - 1) + s) * m;
      const double *Ts = recT + ((size_t)step * (S - 1) + s) * nd;
      /* w = (df/dtd)^T mu per patch */
      for (i = 0; i < n; i++) {
        double Jtd[12];
        rsfo_patch_jacobian(p, i, Ys + (size_t)i * dim, Ts + (size_t)i * nc, NULL, Jtd, NULL,
                            NULL, NULL, NULL, NULL);
        for (c = 0; c < nc; c++) {
          double acc = 0.0;
          for (f = 0; f < dim; f++) acc += Jtd[f * nc + c] * M[s][(size_t)i * dim + f];
          w[i * nc + c] = acc;
        }
      }
      /* z = K^T w */
      for (c = 0; c < nd; c++) z[c] = 0.0;
      for (k = 0; k < nd; k++) {
        const double *kr = p->K + (size_t)k * nd;
        for (c = 0; c < nd; c++) z[c] += kr[c] * w[k];
      }
      /* nu = (df/dx)^T mu + (dv/dx)^T z; parameter gradients */
      for (i = 0; i < n; i++) {
        double Jx[36], Jp[18], Vx[12], Vp[2];
        const double *mu = M[s] + (size_t)i * dim;
        rsfo_patch_jacobian(p, i, Ys + (size_t)i * dim, Ts + (size_t)i * nc, Jx, NULL, Jp, Vx,
                            Vp, NULL, NULL);
        for (int q = 0; q < dim; q++) {
          double acc = 0.0;
          for (f = 0; f < dim; f++) acc += Jx[f * dim + q] * mu[f];
          for (c = 0; c < nc; c++) acc += Vx[c * dim + q] * z[i * nc + c];
          nu[(size_t)i * dim + q] = acc;
        }
        for (f = 0; f < dim; f++) {
          ga[i] += Jp[f * 3 + 0] * mu[f];
          gb[i] += Jp[f * 3 + 1] * mu[f];
          gdc[i] += Jp[f * 3 + 2] * mu[f];
        }
        for (c = 0; c < nc; c++) ga[i] += Vp[c] * z[i * nc + c];
      }
      for (j = 0; j < s; j++)
        for (k = 0; k < m; k++) M[j][k] += h * A[s][j] * nu[k];
      for (k = 0; k < m; k++) lam[k] += nu[k];
    }
    /* misfit at state `step` (= stage 0 of this step) */
    for (i = 0; i < n; i++) {
      double Lx[6], La, td0[2] = {0.0, 0.0};
      const double *x = recY + ((size_t)step * (S - 1)) * m + (size_t)i * dim;
      if (recv_of[i] < 0 || weight[step] == 0.0) continue;
      rsfo_patch_jacobian(p, i, x, td0, NULL, NULL, NULL, NULL, NULL, Lx, &La);
      {
        double res = trace[step * nrecv + recv_of[i]] - target[step * nrecv + recv_of[i]];
        for (f = 0; f < dim; f++) lam[(size_t)i * dim + f] += weight[step] * res * Lx[f];
        ga[i] += weight[step] * res * La;
      }
    }
  }
  for (s = 0; s < 7; s++) free(M[s]);
  free(recY); free(recT); free(trace); free(Xe); free(lam); free(w); free(z); free(nu);
  free(recv_of);
  return J;
}
