// Streaming row-block GEMV core for the dense operator.
//
// A 256-thread workgroup (four wave64s) owns R consecutive rows.  The
// column range is cut into 64-lane chunks of 16-byte vector loads; the four
// waves take contiguous quarters of the chunks, so each wave streams long
// contiguous runs of R rows and reuses every loaded slice of x R times.
// Lane partials are reduced with cross-lane shuffles in FP64, then across
// the four waves through LDS.  Storage formats:
//   double    FP64 entries, FP64 x, FP64 accumulation
//   float     FP32 entries, FP32 x, FP32 lane accumulation
//   bf16x2_t  bf16 hi (upper half) + bf16 lo (lower half) per entry, FP32 x
//   bf16_t    bf16 hi only, FP32 x
// Reads run past the last column up to a whole chunk: callers keep x
// zero-padded to a chunk multiple and allocate that much slack after K.
#pragma once
#include <hip/hip_runtime.h>

#include "rsfg_common.h"

namespace rsfg {

// native clang vector types (the nontemporal builtin needs them)
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <typename TK>
struct KTraits;
template <>
struct KTraits<double> {
  using VT = double;
  using Acc = double;
  static constexpr int VEC = 2;
};
template <>
struct KTraits<float> {
  using VT = float;
  using Acc = float;
  static constexpr int VEC = 4;
};
template <>
struct KTraits<bf16x2_t> {
  using VT = float;
  using Acc = float;
  static constexpr int VEC = 4;
};
template <>
struct KTraits<bf16_t> {
  using VT = float;
  using Acc = float;
  static constexpr int VEC = 8;
};

template <typename TK>
constexpr int chunk_elems() {
  return 64 * KTraits<TK>::VEC;
}

__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }

// acc[r] += K_row_r[col .. col+VEC) . x[col .. col+VEC)
template <typename TK, int R>
struct ChunkFma;

template <int R>
struct ChunkFma<double, R> {
  __device__ __forceinline__ static void run(const double* const* rp, int col,
                                             const double* __restrict__ x,
                                             double* acc) {
    const v2d xv = *reinterpret_cast<const v2d*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v2d kv = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(rp[r] + col));
      acc[r] = fma(kv.x, xv.x, acc[r]);
      acc[r] = fma(kv.y, xv.y, acc[r]);
    }
  }
};

template <int R>
struct ChunkFma<float, R> {
  __device__ __forceinline__ static void run(const float* const* rp, int col,
                                             const float* __restrict__ x,
                                             float* acc) {
    const v4f xv = *reinterpret_cast<const v4f*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4f kv = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(rp[r] + col));
      acc[r] = fmaf(kv.x, xv.x, acc[r]);
      acc[r] = fmaf(kv.y, xv.y, acc[r]);
      acc[r] = fmaf(kv.z, xv.z, acc[r]);
      acc[r] = fmaf(kv.w, xv.w, acc[r]);
    }
  }
};

template <int R>
struct ChunkFma<bf16x2_t, R> {
  __device__ __forceinline__ static void run(const bf16x2_t* const* rp, int col,
                                             const float* __restrict__ x,
                                             float* acc) {
    const v4f xv = *reinterpret_cast<const v4f*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u kv = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(rp[r] + col));
      acc[r] = fmaf(bf_hi(kv.x), xv.x, acc[r]);
      acc[r] = fmaf(bf_lo(kv.x), xv.x, acc[r]);
      acc[r] = fmaf(bf_hi(kv.y), xv.y, acc[r]);
      acc[r] = fmaf(bf_lo(kv.y), xv.y, acc[r]);
      acc[r] = fmaf(bf_hi(kv.z), xv.z, acc[r]);
      acc[r] = fmaf(bf_lo(kv.z), xv.z, acc[r]);
      acc[r] = fmaf(bf_hi(kv.w), xv.w, acc[r]);
      acc[r] = fmaf(bf_lo(kv.w), xv.w, acc[r]);
    }
  }
};

template <int R>
struct ChunkFma<bf16_t, R> {
  __device__ __forceinline__ static void run(const bf16_t* const* rp, int col,
                                             const float* __restrict__ x,
                                             float* acc) {
    const v4f x0 = *reinterpret_cast<const v4f*>(x + col);
    const v4f x1 = *reinterpret_cast<const v4f*>(x + col + 4);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u kv = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(rp[r] + col));
      acc[r] = fmaf(bf_lo(kv.x), x0.x, acc[r]);
      acc[r] = fmaf(bf_hi(kv.x), x0.y, acc[r]);
      acc[r] = fmaf(bf_lo(kv.y), x0.z, acc[r]);
      acc[r] = fmaf(bf_hi(kv.y), x0.w, acc[r]);
      acc[r] = fmaf(bf_lo(kv.z), x1.x, acc[r]);
      acc[r] = fmaf(bf_hi(kv.z), x1.y, acc[r]);
      acc[r] = fmaf(bf_lo(kv.w), x1.z, acc[r]);
      acc[r] = fmaf(bf_hi(kv.w), x1.w, acc[r]);
    }
  }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Row sums of rows [row0, row0 + R) into rowsum[0..R) (LDS, all threads
// see them after return).  red must hold 4 * R doubles of LDS.
template <typename TK, int R>
__device__ __forceinline__ void gemv_rows(const TK* __restrict__ K, long ld,
                                          int row0, int nrows, int cols,
                                          const typename KTraits<TK>::VT* __restrict__ x,
                                          double* rowsum, double* red) {
  using Acc = typename KTraits<TK>::Acc;
  constexpr int CH = chunk_elems<TK>();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nch = (cols + CH - 1) / CH;
  const int per = (nch + 3) >> 2;
  const int c0 = wave * per;
  const int c1 = min(nch, c0 + per);
  const TK* rp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = min(row0 + r, nrows - 1);
    rp[r] = K + (size_t)row * ld;
  }
  Acc acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = Acc(0);
  int c = c0;
  for (; c + 1 < c1; c += 2) {
    ChunkFma<TK, R>::run(rp, c * CH + lane * KTraits<TK>::VEC, x, acc);
    ChunkFma<TK, R>::run(rp, (c + 1) * CH + lane * KTraits<TK>::VEC, x, acc);
  }
  if (c < c1) ChunkFma<TK, R>::run(rp, c * CH + lane * KTraits<TK>::VEC, x, acc);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double s = wave_sum(static_cast<double>(acc[r]));
    if (lane == 0) red[wave * R + r] = s;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    rowsum[r] = (red[r] + red[R + r]) + (red[2 * R + r] + red[3 * R + r]);
  }
  __syncthreads();
}

}  // namespace rsfg
