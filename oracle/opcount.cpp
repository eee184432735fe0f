// Floating-point operation count of one (receiver, source) pair of the
// oracle's Okada path (two DC3D calls, strike and dip slip, half space or
// full space, Hooke's law, rotation back to the global frame, traction on
// the receiver plane and its three projections).  The count is F_pair, the
// constant that converts assembly pairs/s into an algorithmic FLOP rate.
//
// TEST INFRASTRUCTURE ONLY: compiles rsfo_okada.c as C++ with a scalar type
// that counts +, -, *, /, sqrt, log, atan and sin/cos (one each; negation,
// fabs and comparisons are free).  Per-patch work (frames, the patch's own
// sin/cos of strike) is not part of a pair and is not counted.
//
//   make -C oracle opcount && oracle/opcount
#include <cmath>
#include <cstdio>
#include <cstring>

struct OpCount {
  double v;
  OpCount() = default;
  OpCount(double x) : v(x) {}
};

static long n_add, n_mul, n_div, n_sqrt, n_log, n_atan, n_sincos;

static inline OpCount operator+(OpCount a, OpCount b) { ++n_add; return a.v + b.v; }
static inline OpCount operator-(OpCount a, OpCount b) { ++n_add; return a.v - b.v; }
static inline OpCount operator*(OpCount a, OpCount b) { ++n_mul; return a.v * b.v; }
static inline OpCount operator/(OpCount a, OpCount b) { ++n_div; return a.v / b.v; }
static inline OpCount operator-(OpCount a) { return -a.v; }
static inline OpCount& operator+=(OpCount& a, OpCount b) { a = a + b; return a; }
static inline OpCount& operator-=(OpCount& a, OpCount b) { a = a - b; return a; }
static inline bool operator<(OpCount a, OpCount b) { return a.v < b.v; }
static inline bool operator>(OpCount a, OpCount b) { return a.v > b.v; }
static inline bool operator<=(OpCount a, OpCount b) { return a.v <= b.v; }
static inline bool operator>=(OpCount a, OpCount b) { return a.v >= b.v; }
static inline bool operator==(OpCount a, OpCount b) { return a.v == b.v; }
static inline bool operator!=(OpCount a, OpCount b) { return a.v != b.v; }
static inline OpCount fabs(OpCount a) { return std::fabs(a.v); }
static inline OpCount sqrt(OpCount a) { ++n_sqrt; return std::sqrt(a.v); }
static inline OpCount log(OpCount a) { ++n_log; return std::log(a.v); }
static inline OpCount atan(OpCount a) { ++n_atan; return std::atan(a.v); }

#define OKD_REAL OpCount
#define OKD_TO_DOUBLE(x) ((x).v)
#define OKD_NOTE_SINCOS() (++n_sincos, ++n_mul) /* sincos(dip * DEG2RAD) */
#include "rsfo_okada.c"

static void reset() { n_add = n_mul = n_div = n_sqrt = n_log = n_atan = n_sincos = 0; }
static long total() { return n_add + n_mul + n_div + n_sqrt + n_log + n_atan + n_sincos; }

// one pair: receiver 1 km along strike and 2 km up-dip of the source on
// the same plane, or on a parallel plane when off_plane (generic branches:
// off-edge, no negative extension, q != 0)
static long count_pair(float dip, int full_space, int off_plane, double* check) {
  double fr_r[13], fr_s[13], out[2][3];
  int iret[2];
  rsfo_patch_frame(0.0f, dip, fr_r);
  rsfo_patch_frame(0.0f, dip, fr_s);
  const double xs[3] = {0.0, 0.0, -12000.0};
  const double xr[3] = {off_plane ? 700.0 : 0.0, 1000.0, -10000.0};
  reset();
  rsfo_traction_pair(xr, fr_r, xs, fr_s, 500.0, 500.0, full_space, out, iret);
  const long t = total();
  std::printf("dip %5.1f %s %-10s  add/sub %4ld mul %4ld div %3ld sqrt %3ld log %3ld "
              "atan %3ld sincos %2ld  F_pair %5ld   (iret %d %d)\n",
              (double)dip, full_space ? "full" : "half", off_plane ? "off-plane" : "coplanar",
              n_add, n_mul, n_div, n_sqrt, n_log, n_atan, n_sincos, t, iret[0], iret[1]);
  if (check) *check = out[0][0];
  return t;
}

int main() {
  double a, b;
  count_pair(90.0f, 0, 0, &a);
  count_pair(90.0f, 0, 1, &b);
  count_pair(90.0f, 1, 0, nullptr);
  count_pair(60.0f, 0, 1, nullptr);
  count_pair(60.0f, 1, 1, nullptr);
  // the counting type must not change the arithmetic
  double fr[13], out[2][3];
  int ir[2];
  rsfo_patch_frame(0.0f, 90.0f, fr);
  const double xs[3] = {0.0, 0.0, -12000.0}, xr[3] = {0.0, 1000.0, -10000.0};
  rsfo_traction_pair(xr, fr, xs, fr, 500.0, 500.0, 0, out, ir);
  std::printf("check: counted traction %.17g  plain %.17g\n", a, out[0][0]);
  return 0;
}
