// Floating-point operation count of one (receiver, source) pair of the
// oracle's Nikkhoo-Walter triangle path: two half-space triangle stress
// evaluations (unit strike and unit dip slip: main dislocation, image
// dislocation and the free-surface correction of the three edges), Hooke's
// law, and the traction on the receiver plane with its three projections.
// The count is F_pair of triangle assembly, which converts pairs/s into an
// algorithmic FLOP rate.
//
// TEST INFRASTRUCTURE ONLY: compiles rsfo_tri.c (and the generated
// displacement gradients it includes) as C++ with every double replaced by
// a scalar type that counts +, -, *, /, sqrt, log, atan, acos, pow and
// sin/cos (one each; negation, fabs and comparisons are free).  The
// per-triangle work that does not depend on the receiver (TDCS frame, edge
// angles) is done inside the pair by the oracle, as by the reference, and
// is counted.
//
//   make -C oracle opcount_tri && oracle/opcount_tri
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <stdio.h>
#include <cmath>

struct OpCount {
  double v;
  OpCount() = default;
  OpCount(double x) : v(x) {}
};

static long n_add, n_mul, n_div, n_sqrt, n_trans;

static inline OpCount operator+(OpCount a, OpCount b) { ++n_add; return a.v + b.v; }
static inline OpCount operator-(OpCount a, OpCount b) { ++n_add; return a.v - b.v; }
static inline OpCount operator*(OpCount a, OpCount b) { ++n_mul; return a.v * b.v; }
static inline OpCount operator/(OpCount a, OpCount b) { ++n_div; return a.v / b.v; }
static inline OpCount operator-(OpCount a) { return -a.v; }
static inline OpCount& operator+=(OpCount& a, OpCount b) { a = a + b; return a; }
static inline OpCount& operator-=(OpCount& a, OpCount b) { a = a - b; return a; }
static inline OpCount& operator*=(OpCount& a, OpCount b) { a = a * b; return a; }
static inline OpCount& operator/=(OpCount& a, OpCount b) { a = a / b; return a; }
static inline bool operator<(OpCount a, OpCount b) { return a.v < b.v; }
static inline bool operator>(OpCount a, OpCount b) { return a.v > b.v; }
static inline bool operator<=(OpCount a, OpCount b) { return a.v <= b.v; }
static inline bool operator>=(OpCount a, OpCount b) { return a.v >= b.v; }
static inline bool operator==(OpCount a, OpCount b) { return a.v == b.v; }
static inline bool operator!=(OpCount a, OpCount b) { return a.v != b.v; }
static inline OpCount fabs(OpCount a) { return std::fabs(a.v); }
static inline OpCount sqrt(OpCount a) { ++n_sqrt; return std::sqrt(a.v); }
static inline OpCount log(OpCount a) { ++n_trans; return std::log(a.v); }
static inline OpCount atan(OpCount a) { ++n_trans; return std::atan(a.v); }
static inline OpCount acos(OpCount a) { ++n_trans; return std::acos(a.v); }
static inline OpCount pow(OpCount a, OpCount b) { ++n_trans; return std::pow(a.v, b.v); }
static inline void sincos(OpCount a, OpCount* s, OpCount* c) {
  ++n_trans;
  s->v = std::sin(a.v);
  c->v = std::cos(a.v);
}
static inline bool isfinite(OpCount a) { return std::isfinite(a.v); }

#define double OpCount
#include "rsfo_okada.c"
#include "rsfo_tri.c"
#undef double

static void reset() { n_add = n_mul = n_div = n_sqrt = n_trans = 0; }
static long total() { return n_add + n_mul + n_div + n_sqrt + n_trans; }

// ops of the pairs of a mesh of n triangles (all n^2 pairs)
static long count_mesh(int n, const double* vert_in, double* T_out) {
  OpCount xyz[3 * 4], fr[13 * 4], vert[9 * 4], T[6 * 16], l[4], w[4];
  int is_tri[4], iret[2 * 16];
  for (int j = 0; j < n; ++j) {
    is_tri[j] = 1;
    l[j] = w[j] = -1.0;
    for (int k = 0; k < 9; ++k) vert[9 * j + k] = vert_in[9 * j + k];
    OpCount area;
    for (int k = 0; k < 13; ++k) fr[13 * j + k] = 0.0;
    rsfo_tri_frame(vert + 9 * j, vert + 9 * j + 3, vert + 9 * j + 6, fr + 13 * j + 2,
                   fr + 13 * j + 5, fr + 13 * j + 8, &area);
    for (int k = 0; k < 3; ++k)
      xyz[3 * j + k] = (vert_in[9 * j + k] + vert_in[9 * j + 3 + k] + vert_in[9 * j + 6 + k]) / 3.0;
  }
  reset();
  rsfo_mesh_traction_tensor(n, is_tri, xyz, fr, l, w, vert, 0, 1, T, iret);
  const long t = total();
  for (int k = 0; k < 6 * n * n; ++k) T_out[k] = T[k].v;
  for (int k = 0; k < 2 * n * n; ++k)
    if (iret[k]) printf("warning: pair %d singular\n", k / 2);
  return t;
}

int main() {
  // two triangles of one 500 m cell of a gently rough fault plane (x offsets
  // of a few metres), and a third 3 km away along strike and 2 km deeper
  const double v[27] = {
      1.2, -250.0, -10500.0, -0.7, 250.0, -10500.0, 0.4, -250.0, -10000.0,
      -0.7, 250.0, -10500.0, 2.1, 250.0, -10000.0, 0.4, -250.0, -10000.0,
      3.0, 2750.0, -12500.0, -1.5, 3250.0, -12500.0, 0.8, 2750.0, -12000.0};
  double T1[6], T3[6 * 9];
  const long self = count_mesh(1, v, T1);
  const long all3 = count_mesh(3, v, T3);
  const double cross = (all3 - 3.0 * self) / 6.0;
  printf("triangle pairs, half space, strike + dip slip:\n");
  printf("  receiver on the source triangle (its centroid): %ld ops\n", self);
  printf("  3-triangle mesh, 9 pairs: %ld ops, mean over the 6 off-element pairs %.1f\n",
              all3, cross);
  printf("F_pair_tri %ld\n", (long)(cross + 0.5));
  // the counted tractions of the 3-triangle mesh, for comparison with the
  // plain oracle (the counting type must not change the arithmetic)
  printf("tractions");
  for (int k = 0; k < 54; ++k) printf(" %.17g", T3[k]);
  printf("\n");
  return 0;
}
