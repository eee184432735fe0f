// Streaming row-block GEMV core for the dense operator.
//
// One wave64 owns R consecutive rows (the G rows of one patch in the fused
// stage kernel) and streams them over all columns in 64-lane chunks of
// 16-byte nontemporal loads, reusing every loaded slice of x R times.
// Storage formats:
//   double    FP64 entries, FP64 x
//   float     FP32 entries, FP32 x
//   bf16x2_t  bf16 hi (upper half) + bf16 lo (lower half) per entry, FP32 x
//   bf16_t    bf16 hi only, FP32 x
// Every lane keeps its running row sums in FP64.  For the reduced formats
// the products of one 16-byte slice (4 or 8 entries) are summed in FP32 and
// that slice sum is added to the FP64 accumulator, so the accumulation error
// does not grow with the row length; the lane sums are combined in FP64
// with cross-lane shuffles.
// Reads run past the last column up to a whole chunk: callers keep x
// zero-padded to a chunk multiple and allocate that much slack after K.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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
  static constexpr int VEC = 2;
};
template <>
struct KTraits<float> {
  using VT = float;
  static constexpr int VEC = 4;
};
template <>
struct KTraits<bf16x2_t> {
  using VT = float;
  static constexpr int VEC = 4;
};
template <>
struct KTraits<bf16_t> {
  using VT = float;
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
                                             const double* __restrict__ x, double* acc) {
    const v2d xv = *reinterpret_cast<const v2d*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v2d kv = RSFG_KLOAD(reinterpret_cast<const v2d*>(rp[r] + col));
      acc[r] = fma(kv.x, xv.x, acc[r]);
      acc[r] = fma(kv.y, xv.y, acc[r]);
    }
  }
};

template <int R>
struct ChunkFma<float, R> {
  __device__ __forceinline__ static void run(const float* const* rp, int col,
                                             const float* __restrict__ x, double* acc) {
    const v4f xv = *reinterpret_cast<const v4f*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4f kv = RSFG_KLOAD(reinterpret_cast<const v4f*>(rp[r] + col));
      float t = kv.x * xv.x;
      t = fmaf(kv.y, xv.y, t);
      t = fmaf(kv.z, xv.z, t);
      t = fmaf(kv.w, xv.w, t);
      acc[r] += static_cast<double>(t);
    }
  }
};

template <int R>
struct ChunkFma<bf16x2_t, R> {
  __device__ __forceinline__ static void run(const bf16x2_t* const* rp, int col,
                                             const float* __restrict__ x, double* acc) {
    const v4f xv = *reinterpret_cast<const v4f*>(x + col);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u kv = RSFG_KLOAD(reinterpret_cast<const v4u*>(rp[r] + col));
      float t = bf_hi(kv.x) * xv.x;
      t = fmaf(bf_lo(kv.x), xv.x, t);
      t = fmaf(bf_hi(kv.y), xv.y, t);
      t = fmaf(bf_lo(kv.y), xv.y, t);
      t = fmaf(bf_hi(kv.z), xv.z, t);
      t = fmaf(bf_lo(kv.z), xv.z, t);
      t = fmaf(bf_hi(kv.w), xv.w, t);
      t = fmaf(bf_lo(kv.w), xv.w, t);
      acc[r] += static_cast<double>(t);
    }
  }
};

template <int R>
struct ChunkFma<bf16_t, R> {
  __device__ __forceinline__ static void run(const bf16_t* const* rp, int col,
                                             const float* __restrict__ x, double* acc) {
    const v4f x0 = *reinterpret_cast<const v4f*>(x + col);
    const v4f x1 = *reinterpret_cast<const v4f*>(x + col + 4);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u kv = RSFG_KLOAD(reinterpret_cast<const v4u*>(rp[r] + col));
      float t = bf_lo(kv.x) * x0.x;
      t = fmaf(bf_hi(kv.x), x0.y, t);
      t = fmaf(bf_lo(kv.y), x0.z, t);
      t = fmaf(bf_hi(kv.y), x0.w, t);
      t = fmaf(bf_lo(kv.z), x1.x, t);
      t = fmaf(bf_hi(kv.z), x1.y, t);
      t = fmaf(bf_lo(kv.w), x1.z, t);
      t = fmaf(bf_hi(kv.w), x1.w, t);
      acc[r] += static_cast<double>(t);
    }
  }
};

// One 64-lane chunk of R rows held in registers (operator slices k, x
// slices x); fma() does the arithmetic of ChunkFma in the same order (same
// bits) for x slices taken from anywhere (the LDS copy of x below).
template <typename TK>
struct SliceT;
template <>
struct SliceT<double> {
  using KV = v2d;
  using XV = v2d;
  static constexpr int NX = 1;
};
template <>
struct SliceT<float> {
  using KV = v4f;
  using XV = v4f;
  static constexpr int NX = 1;
};
template <>
struct SliceT<bf16x2_t> {
  using KV = v4u;
  using XV = v4f;
  static constexpr int NX = 1;
};
template <>
struct SliceT<bf16_t> {
  using KV = v4u;
  using XV = v4f;
  static constexpr int NX = 2;
};

template <typename TK, int R>
struct Slice {
  using KV = typename SliceT<TK>::KV;
  using XV = typename SliceT<TK>::XV;
  static constexpr int NX = SliceT<TK>::NX;
  KV k[R];
  XV x[NX];
  __device__ __forceinline__ void fma(double* acc) const {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const KV kv = k[r];
      if constexpr (std::is_same<TK, double>::value) {
        acc[r] = ::fma(kv.x, x[0].x, acc[r]);
        acc[r] = ::fma(kv.y, x[0].y, acc[r]);
      } else if constexpr (std::is_same<TK, float>::value) {
        float t = kv.x * x[0].x;
        t = fmaf(kv.y, x[0].y, t);
        t = fmaf(kv.z, x[0].z, t);
        t = fmaf(kv.w, x[0].w, t);
        acc[r] += static_cast<double>(t);
      } else if constexpr (std::is_same<TK, bf16x2_t>::value) {
        float t = bf_hi(kv.x) * x[0].x;
        t = fmaf(bf_lo(kv.x), x[0].x, t);
        t = fmaf(bf_hi(kv.y), x[0].y, t);
        t = fmaf(bf_lo(kv.y), x[0].y, t);
        t = fmaf(bf_hi(kv.z), x[0].z, t);
        t = fmaf(bf_lo(kv.z), x[0].z, t);
        t = fmaf(bf_hi(kv.w), x[0].w, t);
        t = fmaf(bf_lo(kv.w), x[0].w, t);
        acc[r] += static_cast<double>(t);
      } else {
        float t = bf_lo(kv.x) * x[0].x;
        t = fmaf(bf_hi(kv.x), x[0].y, t);
        t = fmaf(bf_lo(kv.y), x[0].z, t);
        t = fmaf(bf_hi(kv.y), x[0].w, t);
        t = fmaf(bf_lo(kv.z), x[1].x, t);
        t = fmaf(bf_hi(kv.z), x[1].y, t);
        t = fmaf(bf_lo(kv.w), x[1].z, t);
        t = fmaf(bf_hi(kv.w), x[1].w, t);
        acc[r] += static_cast<double>(t);
      }
    }
  }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Chunks in flight per loop trip for a wave that streams R rows: enough
// independent 16-byte loads per lane to cover HBM latency, few enough that
// the address streams stay long and sequential.  Measured on gfx950 at
// 8000 and 16000 columns.
template <typename TK, int R>
struct GemvUnroll {
  static constexpr int U = (R == 1) ? 6 : 2;
};
template <int R>
struct GemvUnroll<double, R> {
  static constexpr int U = (R == 1) ? 6 : (R == 2 ? 3 : 2);
};
#ifndef RSFG_BF16_U
#define RSFG_BF16_U 2
#endif
template <int R>
struct GemvUnroll<bf16_t, R> {
  static constexpr int U = (R == 1) ? 4 : RSFG_BF16_U;
};
// Lane partial sums of rows [row0, row0 + R) over all columns, streamed by
// ONE wave (rows past nrows - 1 are clamped to it; the caller discards
// them), added into acc.
template <typename TK, int R>
__device__ __forceinline__ void gemv_acc(const TK* __restrict__ K, long ld, int row0, int nrows,
                                         int cols,
                                         const typename KTraits<TK>::VT* __restrict__ x,
                                         double* acc) {
  constexpr int CH = chunk_elems<TK>();
  constexpr int U = GemvUnroll<TK, R>::U;
  const int lo = (threadIdx.x & 63) * KTraits<TK>::VEC;
  const int nch = (cols + CH - 1) / CH;
  const TK* rp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) rp[r] = K + (size_t)min(row0 + r, nrows - 1) * ld;
  int c = 0;
  for (; c + U <= nch; c += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) ChunkFma<TK, R>::run(rp, (c + u) * CH + lo, x, acc);
  }
  for (; c < nch; ++c) ChunkFma<TK, R>::run(rp, c * CH + lo, x, acc);
}

// Row sums of rows [row0, row0 + R) computed by ONE wave over all columns.
// Every lane holds all R sums on return.  No LDS, no barrier: the waves of
// a workgroup run and finish independently.
template <typename TK, int R>
__device__ __forceinline__ void gemv_wave(const TK* __restrict__ K, long ld, int row0, int nrows,
                                          int cols,
                                          const typename KTraits<TK>::VT* __restrict__ x,
                                          double* out) {
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  gemv_acc<TK, R>(K, ld, row0, nrows, cols, x, acc);
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = wave_sum(acc[r]);
}

// ---- bf16 hi everywhere + bf16 lo in the near-field chunks ("nf")
// The hi parts of all entries stream as the bf16 format does; for the
// nlist column chunks of 512 entries listed for this wave's patch (sorted
// ascending), the lo parts follow in a packed block, chunk k of row r at
// lo[(k R + r) 512].  Entry value = hi + lo where listed, hi elsewhere.
constexpr int kNfChunk = 512;   // = chunk_elems<bf16_t>()

struct bf16nf_t {};   // tag of the format (the hi array is bf16_t)
template <>
struct KTraits<bf16nf_t> {
  using VT = float;
  static constexpr int VEC = 8;
};

template <int R>
__device__ __forceinline__ void gemv_nf_wave(const bf16_t* __restrict__ Khi,
                                             const bf16_t* __restrict__ lo,
                                             const uint16_t* __restrict__ list, int nlist,
                                             long ld, int row0, int nrows, int cols,
                                             const float* __restrict__ x, double* out) {
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  gemv_acc<bf16_t, R>(Khi, ld, row0, nrows, cols, x, acc);
  const int l8 = (threadIdx.x & 63) * 8;
  for (int k = 0; k < nlist; ++k) {
    const int col = (int)list[k] * kNfChunk + l8;
    const v4f x0 = *reinterpret_cast<const v4f*>(x + col);
    const v4f x1 = *reinterpret_cast<const v4f*>(x + col + 4);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const v4u kv = RSFG_KLOAD(
          reinterpret_cast<const v4u*>(lo + ((size_t)k * R + r) * kNfChunk + l8));
      float t = bf_lo(kv.x) * x0.x;
      t = fmaf(bf_hi(kv.x), x0.y, t);
      t = fmaf(bf_lo(kv.y), x0.z, t);
      t = fmaf(bf_hi(kv.y), x0.w, t);
      t = fmaf(bf_lo(kv.z), x1.x, t);
      t = fmaf(bf_hi(kv.z), x1.y, t);
      t = fmaf(bf_lo(kv.w), x1.z, t);
      t = fmaf(bf_hi(kv.w), x1.w, t);
      acc[r] += static_cast<double>(t);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = wave_sum(acc[r]);
}

// ---- x held in LDS
// A workgroup copies x (FP32: the reduced formats) into LDS once per launch
// and its waves stream their rows against that copy, so the slices of x no
// longer travel from L2 with every row pair (for two-byte entries they were
// as many bytes as the operator).  Chunks of the 8-entry formats are stored
// lane-interleaved: lane l's entries 8l .. 8l+3 at 4l and 8l+4 .. 8l+7 at
// 256 + 4l, so both 16-byte LDS reads of a wave are contiguous.  The
// arithmetic is that of ChunkFma, in the same order.
constexpr int kPW = 8;       // waves per workgroup of the persistent kernels (two per SIMD)
constexpr int kXL = 36864;   // floats of x in LDS: 144 KiB of the CU's 160 KiB

// Persistent launch over ncu CUs for `units` work units: upw consecutive
// units per wave and nwg workgroups of kPW waves.  False (no launch) when
// a wave would get more than max_upw units.
inline bool persist_split(int ncu, long units, int max_upw, int* nwg, int* upw) {
  if (ncu <= 0 || units <= 0) return false;
  const long waves = (long)ncu * kPW;
  const long q = (units + waves - 1) / waves;
  if (q > max_upw) return false;
  *upw = (int)q;
  *nwg = (int)((units + kPW * q - 1) / (kPW * q));
  return true;
}

template <typename TK>
constexpr int lds_x_floats(int cols) {
  return (cols + chunk_elems<TK>() - 1) / chunk_elems<TK>() * chunk_elems<TK>();
}

// Column blocks: when x is longer than the LDS copy holds, the persistent
// kernels take the columns in blocks of lds_block_chunks chunks, each copied
// into LDS in turn, and add the blocks' row sums in block order (one block,
// and the sums of the unblocked kernels, when x fits).
template <typename TK>
constexpr int lds_block_chunks() {
  return kXL / chunk_elems<TK>();
}

template <typename TK>
constexpr int lds_blocks(int cols) {
  return (lds_x_floats<TK>(cols) / chunk_elems<TK>() + lds_block_chunks<TK>() - 1) /
         lds_block_chunks<TK>();
}

// chunks [c0, c1) of x into LDS, chunk c at xs + (c - c0) CH
template <typename TK>
__device__ __forceinline__ void x_to_lds(const float* __restrict__ x, int c0, int c1, float* xs) {
  static_assert(std::is_same<typename KTraits<TK>::VT, float>::value, "FP32 x only");
  constexpr int CH = chunk_elems<TK>();
  const float* xb = x + (size_t)c0 * CH;
  const int nslot = (c1 - c0) * CH / KTraits<TK>::VEC;   // (chunk, lane) pairs
  for (int i = threadIdx.x; i < nslot; i += blockDim.x) {
    if constexpr (KTraits<TK>::VEC == 8) {
      const int c = i >> 6, l = i & 63;
      const v4f lo4 = *reinterpret_cast<const v4f*>(xb + c * CH + 8 * l);
      const v4f hi4 = *reinterpret_cast<const v4f*>(xb + c * CH + 8 * l + 4);
      *reinterpret_cast<v4f*>(xs + c * CH + 4 * l) = lo4;
      *reinterpret_cast<v4f*>(xs + c * CH + 256 + 4 * l) = hi4;
    } else {
      *reinterpret_cast<v4f*>(xs + 4 * i) = *reinterpret_cast<const v4f*>(xb + 4 * i);
    }
  }
}

// x slice of chunk c for this lane from the LDS copy
template <typename TK, int R>
__device__ __forceinline__ void lds_x_slice(const float* xs, int c, Slice<TK, R>& sl) {
  constexpr int CH = chunk_elems<TK>();
  const int l4 = (threadIdx.x & 63) * 4;
  if constexpr (KTraits<TK>::VEC == 8) {
    sl.x[0] = *reinterpret_cast<const v4f*>(xs + c * CH + l4);
    sl.x[1] = *reinterpret_cast<const v4f*>(xs + c * CH + 256 + l4);
  } else {
    sl.x[0] = *reinterpret_cast<const v4f*>(xs + c * CH + l4);
  }
}

// Lane partial sums of rows [row0, row0 + R) over the chunks [c0, c1)
// against their LDS copy (chunk c at xs + (c - c0) CH), U chunks per loop
// trip, added into acc (rows past nrows - 1 are clamped to it; the caller
// discards them).
template <typename TK, int R, int U>
__device__ __forceinline__ void gemv_lds_acc(const TK* __restrict__ K, long ld, int row0,
                                             int nrows, int c0, int c1, const float* xs,
                                             double* acc) {
  constexpr int CH = chunk_elems<TK>();
  const int lo = (threadIdx.x & 63) * KTraits<TK>::VEC;
  const TK* rp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) rp[r] = K + (size_t)min(row0 + r, nrows - 1) * ld;
  auto chunk = [&](int c) {
    Slice<TK, R> sl;
#pragma unroll
    for (int r = 0; r < R; ++r)
      sl.k[r] = RSFG_KLOAD(
          reinterpret_cast<const typename Slice<TK, R>::KV*>(rp[r] + (size_t)c * CH + lo));
    lds_x_slice<TK, R>(xs, c - c0, sl);
    sl.fma(acc);
  };
  int c = c0;
  for (; c + U <= c1; c += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) chunk(c + u);
  }
  for (; c < c1; ++c) chunk(c);
}

// The near-field lo chunks of one patch (G rows, nlist chunks listed in
// list, lo block laid out as for gemv_nf_wave) that fall in the column
// block [c0, c1), against its LDS copy.
template <int G>
__device__ __forceinline__ void gemv_nf_lo_lds(const bf16_t* __restrict__ lo,
                                               const uint16_t* __restrict__ list, int nlist,
                                               int c0, int c1, const float* xs, double* acc) {
  const int l8 = (threadIdx.x & 63) * 8;
  // the list is sorted: its entries in [c0, c1) are the run [kb, ke)
  int kb = 0;
  while (kb < nlist && (int)list[kb] < c0) ++kb;
  int ke = kb;
  while (ke < nlist && (int)list[ke] < c1) ++ke;
#pragma unroll 2
  for (int k = kb; k < ke; ++k) {
    Slice<bf16_t, G> sl;
#pragma unroll
    for (int r = 0; r < G; ++r)
      sl.k[r] = RSFG_KLOAD(
          reinterpret_cast<const v4u*>(lo + ((size_t)k * G + r) * kNfChunk + l8));
    lds_x_slice<bf16_t, G>(xs, (int)list[k] - c0, sl);
    sl.fma(acc);
  }
}

}  // namespace rsfg
