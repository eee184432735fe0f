// Device rate-and-state physics: velocity from the regularized law, the
// five state-evolution laws, the quasi-dynamic shear-stress rate and the
// domain check, evaluated per patch in FP64.
//
// Reference behaviour:
//   vel_from_rsf          src/rsf_engine.c:20-30
//   rsf_state_rate        src/rsf_engine.c:44-132
//   rsf_ODE_RHSFunction   src/rsf_engine.c:186-302 (psi first, then sigma,
//                         then tau using both, slip rate = v)
//   rsf_domain_check      src/rsf_engine.c:309-362
#pragma once
#include <hip/hip_runtime.h>

namespace rsfg {

struct FricParams {
  double v0, f0, eta, vmin_state, sato_beta, kt_vc;
  double min_sigma, max_sigma;
  int law, calc_sigma_dot, limit_sigma, ncomp;
};

// per-patch coefficients
struct PatchCoef {
  double a, b, dc;
};

__device__ __forceinline__ double rsf_vel(double tau, double sigma, double psi,
                                          double a, double v0) {
  const double mu = tau / sigma;
  const double st = mu / a;
  const double ef = exp(-psi / a);
  return 2.0 * v0 * ef * sinh(st);
}

// d psi / dt for |v| = vabs_in (the laws only use |v|)
__device__ __forceinline__ double rsf_state_rate(const FricParams& f,
                                                 const PatchCoef& c, double psi,
                                                 double vabs_in) {
  const double b = c.b, D = c.dc;
  if (b == 0.0) return 0.0;
  double vabs = vabs_in;
  switch (f.law) {
    case 2: {  // slip law
      if (vabs < f.vmin_state) vabs = f.vmin_state;
      const double pss = f.f0 - b * log(vabs / f.v0);
      return -(vabs / D) * (psi - pss);
    }
    case 3: {  // PRZ
      const double e = exp((psi - f.f0) / b);
      return (b / (2.0 * D)) * (f.v0 / e - (vabs * vabs / f.v0) * e);
    }
    case 4:
    case 5: {  // Sato / Kato-Tullis
      if (vabs < f.vmin_state) vabs = f.vmin_state;
      const double e = exp((psi - f.f0) / b);
      const double om = (vabs / f.v0) * e;
      const double gate = (f.law == 4) ? exp(-om / f.sato_beta) : exp(-vabs / f.kt_vc);
      const double lom = log(vabs / f.v0) + (psi - f.f0) / b;
      return (b / D) * ((f.v0 / e) * gate - vabs * lom);
    }
    default: {  // aging law
      const double e = exp((f.f0 - psi) / b);
      return (b / D) * (f.v0 * e - vabs);
    }
  }
}

// slip-rate (components) of one patch.  s[]: psi, tau(_s), [tau_d], sigma
template <int NC>
__device__ __forceinline__ void rsf_patch_vel(const FricParams& f, double a,
                                              const double* s, double* v) {
  if (NC == 1) {
    v[0] = rsf_vel(s[1], s[2], s[0], a, f.v0);
  } else {
    const double tm = sqrt(s[1] * s[1] + s[2] * s[2]);
    const double vm = 2.0 * f.v0 * exp(-s[0] / a) * sinh(tm / (s[3] * a));
    if (tm > 0.0) {
      v[0] = vm * s[1] / tm;
      v[1] = vm * s[2] / tm;
    } else {
      v[0] = 0.0;
      v[1] = 0.0;
    }
  }
}

// The right-hand side of one patch is affine in the elastic stressing rates,
// with coefficients that depend on the patch state alone.  rsf_patch_pre
// evaluates that state-only part (slip rate, state rate and the
// radiation-damping terms: every exp, sinh, cosh and division of the law),
// rsf_patch_finish combines it with the stressing rates.  The solver
// evaluates the state-only part when a stage state is formed, one launch
// before its stressing rates exist.  Record layout (PreDim<NC>::v doubles):
//   NC = 1: v, psi_dot, dv/dsigma, dv/dpsi, 1 + eta dv/dtau, sigma
//   NC = 2: v_s, v_d, psi_dot, v_s / a, v_d / a, dvm t_s, dvm t_d, t_s, t_d,
//           1 / (1 + eta g_par), 1 / (1 + eta g_perp), sigma
template <int NC>
struct PreDim {
  static constexpr int v = (NC == 1) ? 6 : 12;
  static constexpr int sig = v - 1;
};

template <int NC>
__device__ __forceinline__ void rsf_patch_pre(const FricParams& f, const PatchCoef& c,
                                              const double* s, double* P) {
  constexpr int iS = NC + 1;  // sigma index
  const double a = c.a, eta = f.eta, sig = s[iS];
  double v[2];
  rsf_patch_vel<NC>(f, a, s, v);
  const double vmag = (NC == 1) ? v[0] : sqrt(v[0] * v[0] + v[1] * v[1]);
  const double fpsi = rsf_state_rate(f, c, s[0], fabs(vmag));
  if (NC == 1) {
    const double mu = s[1] / sig;
    const double st = mu / a;
    const double ef = exp(-s[0] / a);
    const double pre = 2.0 * f.v0 / (a * sig);
    const double cf = cosh(st) * ef;
    const double dvdtau = pre * cf;
    P[0] = v[0];
    P[1] = fpsi;
    P[2] = -pre * cf * mu;     // dv/dsigma
    P[3] = -v[0] / a;          // dv/dpsi
    P[4] = 1.0 + eta * dvdtau;
    P[5] = sig;
  } else {
    // two slip components: dV/dtau = g_par t t^T + g_perp (I - t t^T)
    const double ts = s[1], tdp = s[2];
    const double tm = sqrt(ts * ts + tdp * tdp);
    const double arg = tm / (sig * a);
    const double ef = exp(-s[0] / a);
    const double pre = 2.0 * f.v0 / (a * sig);
    const double gpar = pre * cosh(arg) * ef;
    const double gperp = (arg > 1e-8) ? vmag / tm : pre * ef;
    const double th0 = (tm > 0.0) ? ts / tm : 1.0;
    const double th1 = (tm > 0.0) ? tdp / tm : 0.0;
    const double dvm = -gpar * tm / sig;
    P[0] = v[0];
    P[1] = v[1];
    P[2] = fpsi;
    P[3] = v[0] / a;
    P[4] = v[1] / a;
    P[5] = dvm * th0;
    P[6] = dvm * th1;
    P[7] = th0;
    P[8] = th1;
    P[9] = 1.0 / (1.0 + eta * gpar);
    P[10] = 1.0 / (1.0 + eta * gperp);
    P[11] = sig;
  }
}

// Right-hand side from the state-only record P.  td: elastic stressing K v
// (NC rows); sd: elastic normal rate (In v) or 0; sinc: loading rates (NC
// shear rows) and sinc_n; out: d/dt of the state.
template <int NC>
__device__ __forceinline__ void rsf_patch_finish(const FricParams& f, const double* P,
                                                 const double* td, double sd,
                                                 const double* sinc, double sinc_n,
                                                 double* out) {
  constexpr int iS = NC + 1;
  const double eta = f.eta, sig = P[PreDim<NC>::sig];
  double sdot = f.calc_sigma_dot ? sd + sinc_n : 0.0;
  if (f.limit_sigma && ((sig < f.min_sigma) || (sig > f.max_sigma))) sdot = 0.0;
  out[iS] = sdot;
  if (NC == 1) {
    const double fpsi = P[1];
    out[0] = fpsi;
    double ft = td[0] + sinc[0] - eta * (P[2] * sdot + P[3] * fpsi);
    out[1] = ft / P[4];
    out[3] = P[0];
  } else {
    const double fpsi = P[2], th0 = P[7], th1 = P[8];
    out[0] = fpsi;
    const double r0 = td[0] + sinc[0] - eta * (P[5] * sdot - P[3] * fpsi);
    const double r1 = td[1] + sinc[1] - eta * (P[6] * sdot - P[4] * fpsi);
    const double rp = th0 * r0 + th1 * r1;
    out[1] = th0 * rp * P[9] + (r0 - th0 * rp) * P[10];
    out[2] = th1 * rp * P[9] + (r1 - th1 * rp) * P[10];
    out[4] = P[0];
    out[5] = P[1];
  }
}

// |v| from a state-only record (the same arithmetic as from the state)
template <int NC>
__device__ __forceinline__ double pre_vmag(const double* P) {
  return (NC == 1) ? fabs(P[0]) : sqrt(P[0] * P[0] + P[1] * P[1]);
}

// Right-hand side of one patch (rsf_ODE_RHSFunction, per patch).
template <int NC>
__device__ __forceinline__ void rsf_patch_rhs(const FricParams& f,
                                              const PatchCoef& c,
                                              const double* s, const double* td,
                                              double sd, const double* sinc,
                                              double sinc_n, double* out) {
  double P[PreDim<NC>::v];
  rsf_patch_pre<NC>(f, c, s, P);
  rsf_patch_finish<NC>(f, P, td, sd, sinc, sinc_n, out);
}

// rsf_domain_check for one patch; true if acceptable
template <int NC>
__device__ __forceinline__ bool rsf_patch_ok(const FricParams& f,
                                             const PatchCoef& c,
                                             const double* s) {
  constexpr int iS = NC + 1, dim = 2 * NC + 2;
  const double sig = s[iS];
  if (!(sig > 0.0) || !isfinite(sig) || !(fabs(sig) >= 2.2250738585072014e-308))
    return false;
#pragma unroll
  for (int k = 0; k < dim; ++k)
    if (k != iS && !isfinite(s[k])) return false;
  const double tm = (NC == 1) ? s[1] : sqrt(s[1] * s[1] + s[2] * s[2]);
  if (fabs(tm / (sig * c.a)) > 600.0) return false;
  if (fabs(s[0] / c.a) > 600.0) return false;
  if (f.law != 2 && c.b != 0.0 && fabs((f.f0 - s[0]) / c.b) > 600.0) return false;
  return true;
}

// ---------------------------------------------------------------- IMEX
// Additive split of the scalar-layout system (x = psi, tau, sigma, slip),
// src/rsf_imex.c: explicit G (rsf_IMEX_RHSFunction, :62-115) carries the
// stressing rates,
//   G = (0, (tau_dot + sinc - eta dv/dsigma sigma_dot) / denom, sigma_dot, v),
// implicit Fimpl (rsf_IMEX_IFunction, :123-155, as F = xdot - Fimpl) the
// local state evolution,
//   Fimpl = (S, -eta (dv/dpsi) S / denom, 0, 0),  denom = 1 + eta dv/dtau,
// and G + Fimpl is the right-hand side above.

// state rate with dS/dpsi and dS/d|v| (rsf_state_rate with derivative
// outputs, src/rsf_engine.c:44-132); vin is the signed slip rate
__device__ __forceinline__ double rsf_state_rate_d(const FricParams& f, const PatchCoef& c,
                                                   double psi, double vin, double* dSdpsi,
                                                   double* dSdv) {
  const double b = c.b, D = c.dc;
  if (b == 0.0) {
    *dSdpsi = 0.0;
    *dSdv = 0.0;
    return 0.0;
  }
  double vabs = fabs(vin);
  bool floored = false;
  switch (f.law) {
    case 2: {  // slip law
      if (vabs < f.vmin_state) { vabs = f.vmin_state; floored = true; }
      const double pss = f.f0 - b * log(vabs / f.v0);
      *dSdpsi = -vabs / D;
      *dSdv = floored ? 0.0 : (-(psi - pss) / D - b / D);
      return -(vabs / D) * (psi - pss);
    }
    case 3: {  // PRZ
      const double e = exp((psi - f.f0) / b);
      *dSdpsi = -(1.0 / (2.0 * D)) * (f.v0 / e + (vabs * vabs / f.v0) * e);
      *dSdv = -(b / D) * (vabs / f.v0) * e;
      return (b / (2.0 * D)) * (f.v0 / e - (vabs * vabs / f.v0) * e);
    }
    case 4:
    case 5: {  // Sato / Kato-Tullis
      if (vabs < f.vmin_state) { vabs = f.vmin_state; floored = true; }
      const double e = exp((psi - f.f0) / b);
      const double om = (vabs / f.v0) * e;
      const double gate = (f.law == 4) ? exp(-om / f.sato_beta) : exp(-vabs / f.kt_vc);
      const double lom = log(vabs / f.v0) + (psi - f.f0) / b;
      *dSdpsi = (b / D) * ((f.v0 / e) * (((f.law == 4) ? (-gate * om / (b * f.sato_beta)) : 0.0) -
                                         gate / b) - vabs / b);
      *dSdv = floored ? 0.0
                      : (b / D) * ((f.v0 / e) * ((f.law == 4) ? (-gate * e / (f.v0 * f.sato_beta))
                                                              : (-gate / f.kt_vc)) - lom - 1.0);
      return (b / D) * ((f.v0 / e) * gate - vabs * lom);
    }
    default: {  // aging law
      const double e = exp((f.f0 - psi) / b);
      *dSdpsi = -(f.v0 / D) * e;
      *dSdv = -b / D;
      return (b / D) * (f.v0 * e - vabs);
    }
  }
}

// explicit part G of one cell from its elastic stressing rates td (shear)
// and sd (normal, used when sigma_dot is computed)
__device__ __forceinline__ void imex_explicit(const FricParams& f, const PatchCoef& c,
                                              const double* x, double td, double sd,
                                              double sinc, double sinc_n, double* g) {
  const double a = c.a, eta = f.eta, sig = x[2];
  double sdot = f.calc_sigma_dot ? sd + sinc_n : 0.0;
  if (f.limit_sigma && ((sig < f.min_sigma) || (sig > f.max_sigma))) sdot = 0.0;
  const double mu = x[1] / sig, st = mu / a, ef = exp(-x[0] / a);
  const double pre = 2.0 * f.v0 / (a * sig), cf = cosh(st) * ef;
  const double dvdtau = pre * cf, dvdsig = -pre * cf * mu, denom = 1.0 + eta * dvdtau;
  g[0] = 0.0;
  g[2] = sdot;
  g[1] = (td + sinc - eta * (dvdsig * sdot)) / denom;
  g[3] = rsf_vel(x[1], x[2], x[0], a, f.v0);
}

// implicit part Fimpl of one cell at its state x (the IFunction's terms)
__device__ __forceinline__ void imex_implicit(const FricParams& f, const PatchCoef& c,
                                              const double* x, double* fi) {
  const double a = c.a, eta = f.eta, sig = x[2];
  const double mu = x[1] / sig, st = mu / a, ef = exp(-x[0] / a);
  const double pre = 2.0 * f.v0 / (a * sig), cf = cosh(st) * ef;
  const double denom = 1.0 + eta * (pre * cf);
  const double v = rsf_vel(x[1], x[2], x[0], a, f.v0);
  const double S = rsf_state_rate(f, c, x[0], fabs(v));
  fi[0] = S;
  fi[1] = -(eta * (-v / a) * S) / denom;
  fi[2] = 0.0;
  fi[3] = 0.0;
}

// implicit part (Fimpl_psi, Fimpl_tau) of one cell at (psi, tau; sigma) and,
// with jac, its derivatives with respect to (psi, tau): the psi and tau rows
// of the exact block Jacobian of rsf_IMEX_IJacobian (src/rsf_imex.c:162-237,
// which is shift I minus this block), including the derivatives of the
// radiation-damping prefactor Q = eta (dv/dpsi) / denom
struct ImexLocal {
  double fpsi, ftau, j00, j01, j10, j11;
};

__device__ __forceinline__ void imex_local(const FricParams& f, const PatchCoef& c, double psi,
                                           double tau, double sig, bool jac, ImexLocal& o) {
  const double a = c.a, E = f.eta;
  const double mu = tau / sig, st = mu / a, ef = exp(-psi / a);
  const double v = 2.0 * f.v0 * ef * sinh(st);
  double dSdpsi, dSdv;
  const double S = rsf_state_rate_d(f, c, psi, v, &dSdpsi, &dSdv);
  const double pre = 2.0 * f.v0 / (a * sig), cf = cosh(st) * ef;
  const double dvdtau = pre * cf, dvdpsi = -v / a;
  const double denom = 1.0 + E * dvdtau, Q = E * dvdpsi / denom;
  o.fpsi = S;
  o.ftau = -Q * S;
  if (jac) {
    const double vabs = fabs(v), sg = (v < 0.0) ? -1.0 : 1.0;
    const double Spsi = dSdpsi + dSdv * (-vabs / a);
    const double Stau = dSdv * sg * dvdtau;
    const double d2_psi = -dvdtau / a;
    const double d2_tau = v / (a * a * sig * sig);
    const double k = E / (denom * denom);
    const double Qpsi = k * ((-dvdpsi / a) * denom - dvdpsi * E * d2_psi);
    const double Qtau = k * ((-dvdtau / a) * denom - dvdpsi * E * d2_tau);
    o.j00 = Spsi;
    o.j01 = Stau;
    o.j10 = -(Q * Spsi + S * Qpsi);
    o.j11 = -(Q * Stau + S * Qtau);
  }
}

// Stage equation of one cell, Y = Z + hg Fimpl(Y).  sigma and slip are
// explicit (Y = Z); (psi, tau) solve R(u) = u - Z - hg Fimpl(u) = 0 by
// Newton's method with the exact 2x2 Jacobian I - hg dFimpl/du and
// backtracking on the weighted residual, converged when the weighted
// residual or the weighted update is at most kImexNewtonTol (weights
// 1 / (atol + rtol |Z|), those of the step controller).  y: initial guess
// (psi, tau) in, stage state out; fi: implicit stage derivative
// (Y - Z) / hg.  Returns the iteration count, or -1 when the cell has no
// converged root at this stage size (the step is then retried smaller).
constexpr double kImexNewtonTol = 1e-6;
constexpr int kImexNewtonMaxIt = 50;

__device__ __forceinline__ int imex_newton(const FricParams& f, const PatchCoef& c,
                                           const double* Z, double hg, double rtol,
                                           const double* atol, double* y, double* fi) {
  const double w0 = 1.0 / (atol[0] + rtol * fabs(Z[0]));
  const double w1 = 1.0 / (atol[1] + rtol * fabs(Z[1]));
  double u0 = y[0], u1 = y[1];
  ImexLocal L;
  imex_local(f, c, u0, u1, Z[2], true, L);
  double r0 = u0 - Z[0] - hg * L.fpsi;
  double r1 = u1 - Z[1] - hg * L.ftau;
  double rn = fmax(fabs(r0) * w0, fabs(r1) * w1);
  int it = 0;
  bool conv = false;
  for (; it < kImexNewtonMaxIt; ++it) {
    if (!isfinite(rn)) return -1;
    if (rn <= kImexNewtonTol) {
      conv = true;
      break;
    }
    const double m00 = 1.0 - hg * L.j00, m01 = -hg * L.j01;
    const double m10 = -hg * L.j10, m11 = 1.0 - hg * L.j11;
    const double det = m00 * m11 - m01 * m10;
    if (!(det != 0.0) || !isfinite(det)) return -1;
    const double d0 = (-r0 * m11 + r1 * m01) / det;
    const double d1 = (-r1 * m00 + r0 * m10) / det;
    double lam = 1.0, t0 = 0.0, t1 = 0.0, rt0 = 0.0, rt1 = 0.0, rtn = 0.0;
    bool ok = false;
    for (int k = 0; k < 12; ++k) {
      t0 = u0 + lam * d0;
      t1 = u1 + lam * d1;
      imex_local(f, c, t0, t1, Z[2], true, L);
      rt0 = t0 - Z[0] - hg * L.fpsi;
      rt1 = t1 - Z[1] - hg * L.ftau;
      rtn = fmax(fabs(rt0) * w0, fabs(rt1) * w1);
      if (isfinite(rtn) && rtn < (1.0 - 1e-4 * lam) * rn) {
        ok = true;
        break;
      }
      lam *= 0.5;
    }
    if (!ok) return -1;
    u0 = t0;
    u1 = t1;
    r0 = rt0;
    r1 = rt1;
    rn = rtn;
    if (fmax(fabs(lam * d0) * w0, fabs(lam * d1) * w1) <= kImexNewtonTol) {
      conv = true;
      ++it;
      break;
    }
  }
  if (!conv) return -1;
  y[0] = u0;
  y[1] = u1;
  y[2] = Z[2];
  y[3] = Z[3];
  fi[0] = (u0 - Z[0]) / hg;
  fi[1] = (u1 - Z[1]) / hg;
  fi[2] = 0.0;
  fi[3] = 0.0;
  return it;
}

}  // namespace rsfg
