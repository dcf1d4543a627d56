// Device-side helpers shared by the index-build kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

namespace pg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kTomb = 0xFFFFFFFFu;
constexpr uint32_t kMaskedTomb = 0x7FFFFFFFu;  // tomb after & 0x7FFFFFFF
constexpr uint32_t kMsb = 0x80000000u;
constexpr int kWave = 64;
constexpr int kStatSlots = 256;  // spread counter atomics over this many rows
constexpr int kStatFields = 16;
constexpr uint64_t kPadKey = ~0ull;

enum StatField {
  ST_PRUNE_DIST = 0,
  ST_PRUNE_PAIRS,
  ST_MERGE_DIST,
  ST_MERGE_PAIRS,
  ST_SEARCH_DIST,
  ST_SEARCH_ITERS,
  ST_SEARCH_CAPPED,
  ST_REV_OVERFLOW,
  ST_SCREEN_FALLBACK,
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Float -> uint32 whose unsigned order equals the float order (no NaNs).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & kMsb) ? ~u : (u | kMsb);
}
__device__ __forceinline__ float from_ordered(uint32_t o) {
  return __uint_as_float((o & kMsb) ? (o & ~kMsb) : ~o);
}
__device__ __forceinline__ uint32_t rotl1(uint32_t x) { return (x << 1) | (x >> 31); }
__device__ __forceinline__ uint32_t rotr1(uint32_t x) { return (x >> 1) | (x << 31); }

// Search-list key: (distance, id & 0x7FFFFFFF, expanded bit) in one u64.
__device__ __forceinline__ uint64_t search_key(float d, uint32_t raw_id) {
  return (uint64_t(ordered_bits(d)) << 32) | rotl1(raw_id);
}
__device__ __forceinline__ uint32_t key_raw_id(uint64_t k) {
  return rotr1(uint32_t(k));
}
__device__ __forceinline__ float key_dist(uint64_t k) {
  return from_ordered(uint32_t(k >> 32));
}

__device__ __forceinline__ bool is_stop_id(uint32_t id, uint32_t n) {
  return id == kTomb || id == kMaskedTomb || id >= n;
}

// Sequential FMA chain over the whole row (Gbuilder.cuh:785-797), negated.
// Loads are issued 8 float4 at a time per operand so that one lane keeps
// 16 row segments in flight; the FMAs stay in index order.
__device__ __forceinline__ float chain_dist(const float* __restrict__ x,
                                            const float* __restrict__ y,
                                            uint32_t dim) {
  float s = 0.0f;
  if ((dim & 3u) == 0) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* y4 = reinterpret_cast<const float4*>(y);
    const uint32_t n4 = dim >> 2;
    uint32_t i = 0;
    for (; i + 8 <= n4; i += 8) {
      float4 a[8], b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        a[u] = x4[i + u];
        b[u] = y4[i + u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        s = __builtin_fmaf(a[u].x, b[u].x, s);
        s = __builtin_fmaf(a[u].y, b[u].y, s);
        s = __builtin_fmaf(a[u].z, b[u].z, s);
        s = __builtin_fmaf(a[u].w, b[u].w, s);
      }
    }
    for (; i < n4; ++i) {
      const float4 a = x4[i], b = y4[i];
      s = __builtin_fmaf(a.x, b.x, s);
      s = __builtin_fmaf(a.y, b.y, s);
      s = __builtin_fmaf(a.z, b.z, s);
      s = __builtin_fmaf(a.w, b.w, s);
    }
  } else {
    for (uint32_t i = 0; i < dim; ++i) s = __builtin_fmaf(x[i], y[i], s);
  }
  return -s;
}

// Lane k's value (k in 0..3) of the caller's quad, by a DPP quad broadcast.
template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
  return __int_as_float(
      __builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, false));
}

// chain_dist(x, y_q, dim) for one row y_q per quad (lanes 4q .. 4q+3), all
// 64 lanes taking part (no lane may be inactive: the operands travel by DPP).
// Lane k of a quad loads float4s 4j + k of its row, JB per batch, all in
// flight together; then every lane of the quad runs the same sequential FMA
// chain, taking float4 f from lane f % 4 by a quad broadcast and x from x_lds
// (one row for the whole wave, in LDS: broadcast reads).  Same FMAs in the
// same order as chain_dist, so the result is bit-equal.  Rows need dim % 4 ==
// 0 (16-byte aligned float4s); other widths take chain_dist per lane.
// Quads with !on return 0.
template <int JB>
__device__ __forceinline__ float quad_chain(const float* x_lds,
                                            const float* __restrict__ y,
                                            bool on, uint32_t dim) {
  if (dim & 3u) return on ? chain_dist(x_lds, y, dim) : 0.0f;
  const uint32_t k = threadIdx.x & 3u;
  const uint32_t n4 = dim >> 2;
  const float4* y4 = reinterpret_cast<const float4*>(y);
  const float4* x4 = reinterpret_cast<const float4*>(x_lds);
  float s = 0.0f;
  for (uint32_t f0 = 0; f0 < n4; f0 += 4u * JB) {
    float4 v[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      const uint32_t f = f0 + 4u * j + k;
      v[j] = (on && f < n4) ? y4[f] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const uint32_t f = f0 + 4u * j + kk;
        if (f < n4) {
          const float4 a = x4[f];
          float4 b;
          if (kk == 0) {
            b.x = quad_bcast<0>(v[j].x); b.y = quad_bcast<0>(v[j].y);
            b.z = quad_bcast<0>(v[j].z); b.w = quad_bcast<0>(v[j].w);
          } else if (kk == 1) {
            b.x = quad_bcast<1>(v[j].x); b.y = quad_bcast<1>(v[j].y);
            b.z = quad_bcast<1>(v[j].z); b.w = quad_bcast<1>(v[j].w);
          } else if (kk == 2) {
            b.x = quad_bcast<2>(v[j].x); b.y = quad_bcast<2>(v[j].y);
            b.z = quad_bcast<2>(v[j].z); b.w = quad_bcast<2>(v[j].w);
          } else {
            b.x = quad_bcast<3>(v[j].x); b.y = quad_bcast<3>(v[j].y);
            b.z = quad_bcast<3>(v[j].z); b.w = quad_bcast<3>(v[j].w);
          }
          s = __builtin_fmaf(a.x, b.x, s);
          s = __builtin_fmaf(a.y, b.y, s);
          s = __builtin_fmaf(a.z, b.z, s);
          s = __builtin_fmaf(a.w, b.w, s);
        }
      }
    }
  }
  return on ? -s : 0.0f;
}

// Value held by lane (lane ^ M), M in {1, 2, 4, 8, 16, 32}, without an LDS
// round trip: DPP quad permutations for 1 and 2, DPP row rotate by 8 for 8
// (within a 16-lane row, rotating by 8 is xor 8), ds_swizzle in bit-mask
// mode for 4 (xor inside 32-lane halves; no address register), and the
// gfx950 lane swaps v_permlane16_swap / v_permlane32_swap for 16 and 32
// (each swap hands one of its two results the partner row's values).
template <uint32_t M>
__device__ __forceinline__ uint32_t xor_lane32(uint32_t v) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32,
                "lane xor mask");
  if constexpr (M == 1) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, false));
  } else if constexpr (M == 2) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x4E, 0xF, 0xF, false));
  } else if constexpr (M == 4) {
    return uint32_t(__builtin_amdgcn_ds_swizzle(int(v), 0x101F));
  } else if constexpr (M == 8) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), 0x128, 0xF, 0xF, false));
  } else if constexpr (M == 16) {
    // rows 0 and 2 read the odd rows from the second result, rows 1 and 3
    // the even rows from the first
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane_id() & 16u) ? p[0] : p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane_id() & 32u) ? p[0] : p[1];
  }
}
// Any 4- or 8-byte value, by its bits.
template <uint32_t M, typename T>
__device__ __forceinline__ T xor_lane(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, xor_lane32<M>(__builtin_bit_cast(uint32_t, v)));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t(xor_lane32<M>(uint32_t(u >> 32))) << 32) |
                       xor_lane32<M>(uint32_t(u));
    return __builtin_bit_cast(T, r);
  }
}
// Same with the mask as a value that is a constant after unrolling.
template <typename T>
__device__ __forceinline__ T xor_lane_dyn(T v, uint32_t m) {
  switch (m) {
    case 1: return xor_lane<1, T>(v);
    case 2: return xor_lane<2, T>(v);
    case 4: return xor_lane<4, T>(v);
    case 8: return xor_lane<8, T>(v);
    case 16: return xor_lane<16, T>(v);
    default: return xor_lane<32, T>(v);
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += xor_lane<32, T>(v);
  v += xor_lane<16, T>(v);
  v += xor_lane<8, T>(v);
  v += xor_lane<4, T>(v);
  v += xor_lane<2, T>(v);
  v += xor_lane<1, T>(v);
  return v;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// Orders the LDS traffic of one wave between its lanes without a
// workgroup barrier: a wave's LDS instructions execute in issue order, so
// it is enough that the compiler keeps its loads and stores on their side
// of this point.  Unlike __syncthreads it waits for nothing (global loads
// stay in flight) and leaves the other waves of the workgroup alone.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier of the kernels whose workgroup is one wave (__syncthreads), or,
// with kWaveSync, of one wave inside a larger workgroup (wave_sync).
template <bool kWaveSync>
__device__ __forceinline__ void lds_sync() {
  if constexpr (kWaveSync) wave_sync();
  else __syncthreads();
}

// Ascending bitonic sort of n (power of two, >= 2) keys in LDS by one wave
// (lds_sync orders the LDS traffic between its stages).
template <typename K, bool kWaveSync = false>
__device__ void bitonic_sort(K* a, uint32_t n) {
  static_assert(kWave == 64, "lanes stride by the wave size");
  const uint32_t lane = lane_id();
  const uint32_t half = n >> 1;
  for (uint32_t k = 2; k <= n; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = lane; t < half; t += kWave) {
        uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        uint32_t p = i | j;
        K x = a[i], y = a[p];
        bool asc = (i & k) == 0;
        if ((x > y) == asc) {
          a[i] = y;
          a[p] = x;
        }
      }
      lds_sync<kWaveSync>();
    }
  }
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, uint32_t m) {
  return xor_lane_dyn(v, m);
}

__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, uint32_t m) {
  return xor_lane_dyn(v, m);
}
__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, uint32_t m) {
  return xor_lane_dyn(v, m);
}

// Ascending bitonic sort (Batcher) of 64*R keys (uint32_t or uint64_t) held
// in registers, element index lane*R + r.  Cross-lane compare-exchanges take
// the partner through xor_lane (DPP / ds_swizzle / lane swaps: no LDS
// traffic and no address registers); in-lane ones swap registers.
template <int R, typename K>
__device__ __forceinline__ void reg_bitonic_sort(K (&v)[R]) {
  const uint32_t lane = lane_id();
  constexpr uint32_t n = 64u * R;
#pragma unroll
  for (uint32_t k = 2; k <= n; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= uint32_t(R)) {
        const uint32_t lj = j / R;
        const bool lower = (lane & lj) == 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t e = lane * R + r;
          const bool asc = (e & k) == 0;
          const K o = shfl_xor_key(v[r], lj);
          const K mn = v[r] < o ? v[r] : o;
          const K mx = v[r] < o ? o : v[r];
          v[r] = (lower == asc) ? mn : mx;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((uint32_t(r) & j) == 0) {
            const uint32_t e = lane * R + r;
            const bool asc = (e & k) == 0;
            const K a = v[r], b = v[r + j];
            if ((a > b) == asc) {
              v[r] = b;
              v[r + j] = a;
            }
          }
        }
      }
    }
  }
}

// Ascending position of this lane's value among the values of lanes
// [0, n) (n <= 64, uniform): #(lanes i < n whose (value, i) is below this
// lane's (value, lane)), so equal values take consecutive positions in lane
// order and the n positions are a permutation of [0, n).  One readlane and
// a compare per value: for short lists this is far cheaper than a 64-wide
// bitonic network.  Every lane must take part; lanes >= n get an unused
// result.
__device__ __forceinline__ uint32_t wave_rank(uint32_t v, uint32_t n) {
  const uint32_t lane = lane_id();
  uint32_t r = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t a = __builtin_amdgcn_readlane(v, i);
    r += (a < v || (a == v && i < lane)) ? 1u : 0u;
  }
  return r;
}
// The same for 64-bit keys that are distinct among lanes [0, n).
__device__ __forceinline__ uint32_t wave_rank_distinct(uint64_t v,
                                                       uint32_t n) {
  const uint32_t hi = uint32_t(v >> 32), lo = uint32_t(v);
  uint32_t r = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t a =
        (uint64_t(__builtin_amdgcn_readlane(hi, i)) << 32) |
        uint32_t(__builtin_amdgcn_readlane(lo, i));
    r += a < v ? 1u : 0u;
  }
  return r;
}

// Number of entries of the sorted LDS array a[0..n) that are < x
// (lower bound) or <= x (upper bound).
__device__ __forceinline__ uint32_t count_less(const uint64_t* a, uint32_t n,
                                               uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t count_less_eq(const uint64_t* a,
                                                  uint32_t n, uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void stat_add(unsigned long long* stats, int field,
                                         unsigned long long v) {
  if (stats != nullptr && v != 0)
    atomicAdd(&stats[(blockIdx.x % kStatSlots) * kStatFields + field], v);
}

// Ascending sort of the n keys a[0..n) in LDS (n a power of two): in
// registers (64*R keys, shuffle exchanges, no barriers between stages) for n
// up to 512, else the LDS bitonic sort.  The caller's last LDS writes must be
// behind a barrier.
template <int R, bool kWaveSync = false, typename K>
__device__ __forceinline__ void reg_sort_lds(K* a) {
  const uint32_t lane = lane_id();
  K v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = a[lane * R + r];
  reg_bitonic_sort<R>(v);
  lds_sync<kWaveSync>();
#pragma unroll
  for (int r = 0; r < R; ++r) a[lane * R + r] = v[r];
  lds_sync<kWaveSync>();
}

template <bool kWaveSync = false, typename K>
__device__ __forceinline__ void sort_list(K* a, uint32_t n) {
  if (n == 64) reg_sort_lds<1, kWaveSync>(a);
  else if (n == 128) reg_sort_lds<2, kWaveSync>(a);
  else if (n == 256) reg_sort_lds<4, kWaveSync>(a);
  else if (n == 512) reg_sort_lds<8, kWaveSync>(a);
  else bitonic_sort<K, kWaveSync>(a, n);
}

// Screened RNG (the BF16 path): the same decisions as rng_select, with most
// pair distances taken from BF16 MFMA Gram tiles.  Candidates are resolved

// Elements e0 .. e0+7 of a query row (zero past dim or when !on): two
// 16-byte loads when the row start is 16-byte aligned and dim % 8 == 0,
// else scalar loads.
__device__ __forceinline__ void load_query_octet(const float* __restrict__ x,
                                                 uint32_t dim, bool on,
                                                 uint32_t e0, float (&v)[8]) {
  if (on && (dim & 7u) == 0 && e0 + 8u <= dim &&
      (reinterpret_cast<uintptr_t>(x) & 15u) == 0) {
    const float4 a = *reinterpret_cast<const float4*>(x + e0);
    const float4 b = *reinterpret_cast<const float4*>(x + e0 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      v[j] = (on && e0 + j < dim) ? x[e0 + j] : 0.0f;
  }
}

}  // namespace pg
