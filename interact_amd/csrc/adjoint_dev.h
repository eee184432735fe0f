// Pointwise rate-and-state physics written for a generic scalar type, for
// the discrete adjoint of the explicit Runge-Kutta steps.
//
// T = double gives the right-hand side of one patch (the same formulas as
// rsf_patch_pre + rsf_patch_finish: rsf_ODE_RHSFunction,
// src/rsf_engine.c:186-302, with rsf_state_rate :44-132 and vel_from_rsf
// :20-30); T = Dual carries one forward-mode tangent through the same
// arithmetic, so a directional derivative of the patch right-hand side (or
// of its slip rate) costs one evaluation.  The adjoint kernels apply the
// transposed local Jacobian to an adjoint vector one input direction at a
// time: d/ds_q, d/dtd_k and d/d(a, b, dc).
#pragma once
#include <hip/hip_runtime.h>

#include "rsf_dev.h"

namespace rsfg {
namespace adj {

struct Dual {
  double v, d;
};

__device__ __forceinline__ Dual operator+(Dual x, Dual y) { return {x.v + y.v, x.d + y.d}; }
__device__ __forceinline__ Dual operator-(Dual x, Dual y) { return {x.v - y.v, x.d - y.d}; }
__device__ __forceinline__ Dual operator*(Dual x, Dual y) { return {x.v * y.v, x.d * y.v + x.v * y.d}; }
__device__ __forceinline__ Dual operator/(Dual x, Dual y) {
  const double q = x.v / y.v;
  return {q, (x.d - q * y.d) / y.v};
}
__device__ __forceinline__ Dual operator-(Dual x) { return {-x.v, -x.d}; }
__device__ __forceinline__ Dual operator+(Dual x, double c) { return {x.v + c, x.d}; }
__device__ __forceinline__ Dual operator+(double c, Dual x) { return {c + x.v, x.d}; }
__device__ __forceinline__ Dual operator-(Dual x, double c) { return {x.v - c, x.d}; }
__device__ __forceinline__ Dual operator-(double c, Dual x) { return {c - x.v, -x.d}; }
__device__ __forceinline__ Dual operator*(Dual x, double c) { return {x.v * c, x.d * c}; }
__device__ __forceinline__ Dual operator*(double c, Dual x) { return {c * x.v, c * x.d}; }
__device__ __forceinline__ Dual operator/(Dual x, double c) { return {x.v / c, x.d / c}; }
__device__ __forceinline__ Dual operator/(double c, Dual x) {
  const double q = c / x.v;
  return {q, -q * x.d / x.v};
}
__device__ __forceinline__ Dual exp(Dual x) {
  const double e = ::exp(x.v);
  return {e, e * x.d};
}
__device__ __forceinline__ Dual log(Dual x) { return {::log(x.v), x.d / x.v}; }
__device__ __forceinline__ Dual sinh(Dual x) { return {::sinh(x.v), ::cosh(x.v) * x.d}; }
__device__ __forceinline__ Dual cosh(Dual x) { return {::cosh(x.v), ::sinh(x.v) * x.d}; }
__device__ __forceinline__ Dual sqrt(Dual x) {
  const double s = ::sqrt(x.v);
  return {s, (s > 0.0) ? x.d / (2.0 * s) : 0.0};
}
__device__ __forceinline__ Dual fabs(Dual x) { return (x.v < 0.0) ? -x : x; }

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
__device__ __forceinline__ double exp(double x) { return ::exp(x); }
__device__ __forceinline__ double log(double x) { return ::log(x); }
__device__ __forceinline__ double sinh(double x) { return ::sinh(x); }
__device__ __forceinline__ double cosh(double x) { return ::cosh(x); }
__device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ double fabs(double x) { return ::fabs(x); }

template <class T>
__device__ __forceinline__ T cst(double c) {
  return T{c};
}
template <>
__device__ __forceinline__ Dual cst<Dual>(double c) {
  return Dual{c, 0.0};
}

// rsf_state_rate for |v| = vabs_in (src/rsf_engine.c:44-132)
template <class T>
__device__ __forceinline__ T state_rate(const FricParams& f, T b, T D, T psi, T vabs_in) {
  if (val(b) == 0.0) return cst<T>(0.0);
  T vabs = vabs_in;
  switch (f.law) {
    case 2: {  // slip law
      if (val(vabs) < f.vmin_state) vabs = cst<T>(f.vmin_state);
      const T pss = f.f0 - b * log(vabs / f.v0);
      return -(vabs / D) * (psi - pss);
    }
    case 3: {  // PRZ
      const T e = exp((psi - f.f0) / b);
      return (b / (2.0 * D)) * (f.v0 / e - (vabs * vabs / f.v0) * e);
    }
    case 4:
    case 5: {  // Sato / Kato-Tullis
      if (val(vabs) < f.vmin_state) vabs = cst<T>(f.vmin_state);
      const T e = exp((psi - f.f0) / b);
      const T om = (vabs / f.v0) * e;
      const T gate = (f.law == 4) ? exp(-om / f.sato_beta) : exp(-vabs / f.kt_vc);
      const T lom = log(vabs / f.v0) + (psi - f.f0) / b;
      return (b / D) * ((f.v0 / e) * gate - vabs * lom);
    }
    default: {  // aging law
      const T e = exp((f.f0 - psi) / b);
      return (b / D) * (f.v0 * e - vabs);
    }
  }
}

// slip rate (components) of one patch; s: psi, tau(_s), [tau_d], sigma, ...
template <class T, int NC>
__device__ __forceinline__ void vel(const FricParams& f, T a, const T* s, T* v) {
  if (NC == 1) {
    v[0] = 2.0 * f.v0 * exp(-s[0] / a) * sinh((s[1] / s[2]) / a);
  } else {
    const T tm = sqrt(s[1] * s[1] + s[2] * s[2]);
    const T vm = 2.0 * f.v0 * exp(-s[0] / a) * sinh(tm / (s[3] * a));
    if (val(tm) > 0.0) {
      v[0] = vm * s[1] / tm;
      v[1] = vm * s[2] / tm;
    } else {
      v[0] = cst<T>(0.0);
      v[1] = cst<T>(0.0);
    }
  }
}

// right-hand side of one patch from its state s and elastic stressing
// rates td (NC shear rows; no normal rows), loading sinc[NC], sinc_n
template <class T, int NC>
__device__ __forceinline__ void rhs(const FricParams& f, T a, T b, T dc, const T* s, const T* td,
                                    const double* sinc, double sinc_n, T* out) {
  constexpr int iS = NC + 1;
  const double eta = f.eta;
  const T sig = s[iS];
  T v[2];
  vel<T, NC>(f, a, s, v);
  const T vmag = (NC == 1) ? v[0] : sqrt(v[0] * v[0] + v[1] * v[1]);
  const T fpsi = state_rate<T>(f, b, dc, s[0], fabs(vmag));
  double sdot = f.calc_sigma_dot ? sinc_n : 0.0;
  if (f.limit_sigma && ((val(sig) < f.min_sigma) || (val(sig) > f.max_sigma))) sdot = 0.0;
  out[0] = fpsi;
  out[iS] = cst<T>(sdot);
  if (NC == 1) {
    const T mu = s[1] / sig;
    const T st = mu / a;
    const T ef = exp(-s[0] / a);
    const T pre = 2.0 * f.v0 / (a * sig);
    const T cf = cosh(st) * ef;
    const T dvdtau = pre * cf;
    const T dvdsig = -(pre * cf * mu);
    const T dvdpsi = -(v[0] / a);
    out[1] = (td[0] + sinc[0] - eta * (dvdsig * sdot + dvdpsi * fpsi)) / (1.0 + eta * dvdtau);
    out[3] = v[0];
  } else {
    const T ts = s[1], tdp = s[2];
    const T tm = sqrt(ts * ts + tdp * tdp);
    const T arg = tm / (sig * a);
    const T ef = exp(-s[0] / a);
    const T pre = 2.0 * f.v0 / (a * sig);
    const T gpar = pre * cosh(arg) * ef;
    const T gperp = (val(arg) > 1e-8) ? vmag / tm : pre * ef;
    const T th0 = (val(tm) > 0.0) ? ts / tm : cst<T>(1.0);
    const T th1 = (val(tm) > 0.0) ? tdp / tm : cst<T>(0.0);
    const T dvm = -(gpar * tm / sig);
    const T r0 = td[0] + sinc[0] - eta * (dvm * th0 * sdot - (v[0] / a) * fpsi);
    const T r1 = td[1] + sinc[1] - eta * (dvm * th1 * sdot - (v[1] / a) * fpsi);
    const T rp = th0 * r0 + th1 * r1;
    const T dpar = 1.0 + eta * gpar, dperp = 1.0 + eta * gperp;
    out[1] = th0 * rp / dpar + (r0 - th0 * rp) / dperp;
    out[2] = th1 * rp / dpar + (r1 - th1 * rp) / dperp;
    out[4] = v[0];
    out[5] = v[1];
  }
}

}  // namespace adj
}  // namespace rsfg
