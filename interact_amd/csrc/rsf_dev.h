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

// Right-hand side of one patch.  s: state (psi, tau.., sigma, slip..);
// td: elastic stressing K v (NC rows); sd: elastic normal rate (In v) or 0;
// sinc: loading rates (NC shear rows) and sinc_n; out: d/dt of the state.
template <int NC>
__device__ __forceinline__ void rsf_patch_rhs(const FricParams& f,
                                              const PatchCoef& c,
                                              const double* s, const double* td,
                                              double sd, const double* sinc,
                                              double sinc_n, double* out) {
  constexpr int iS = NC + 1;  // sigma index
  const double a = c.a, eta = f.eta, sig = s[iS];
  double v[2];
  rsf_patch_vel<NC>(f, a, s, v);
  const double vmag = (NC == 1) ? v[0] : sqrt(v[0] * v[0] + v[1] * v[1]);
  const double fpsi = rsf_state_rate(f, c, s[0], fabs(vmag));
  double sdot = f.calc_sigma_dot ? sd + sinc_n : 0.0;
  if (f.limit_sigma && ((sig < f.min_sigma) || (sig > f.max_sigma))) sdot = 0.0;
  out[0] = fpsi;
  out[iS] = sdot;
  if (NC == 1) {
    const double mu = s[1] / sig;
    const double st = mu / a;
    const double ef = exp(-s[0] / a);
    const double pre = 2.0 * f.v0 / (a * sig);
    const double cf = cosh(st) * ef;
    const double dvdtau = pre * cf;
    const double dvdsig = -pre * cf * mu;
    const double dvdpsi = -v[0] / a;
    double ft = td[0] + sinc[0] - eta * (dvdsig * sdot + dvdpsi * fpsi);
    out[1] = ft / (1.0 + eta * dvdtau);
    out[3] = v[0];
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
    const double r0 = td[0] + sinc[0] - eta * (dvm * th0 * sdot - v[0] / a * fpsi);
    const double r1 = td[1] + sinc[1] - eta * (dvm * th1 * sdot - v[1] / a * fpsi);
    const double rp = th0 * r0 + th1 * r1;
    const double ipar = 1.0 / (1.0 + eta * gpar), iperp = 1.0 / (1.0 + eta * gperp);
    out[1] = th0 * rp * ipar + (r0 - th0 * rp) * iperp;
    out[2] = th1 * rp * ipar + (r1 - th1 * rp) * iperp;
    out[4] = v[0];
    out[5] = v[1];
  }
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

}  // namespace rsfg
