// Exact top-k by inner product (the gt the index build reads).
//
// The reference reads its gt from the dataset's ground-truth file, padded
// to 128 columns (tests/test_t2iend.cu:735-760, transform_gt_to_128.ipynb);
// for the synthetic workloads this computes it on the GPU for any size.
// Definition: the k base ids with the largest inner product q.x, scored by
// the FP32 sequential FMA chain (the prune's distance order), ties to the
// lower id.
//
// Screen: BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate) over the
// BF16 copies of queries and base.  For every pair the screened s obeys
// |s - chain(q, x)| <= e, the per-pair bound of the screened prune: with
// n = |x| and r = |x - bf16(x)| per row and h = n + r,
//   e = 1.01 (h_q r_x + r_q h_x + r_q r_x + g_a h_q h_x + g_c n_q n_x)
// (the BF16 rounding of both operands by Cauchy-Schwarz, the MFMA's FP32
// accumulation over dpad terms, the chain's own error over dim terms).
// Each query keeps a candidate buffer and a threshold t, a lower bound of
// its true k-th largest score: a pair is kept iff s + e >= t, so the
// buffer always holds every true top-k id.  When a buffer fills, its wave
// sorts it by s - e, sets t to the k-th largest s - e (k kept pairs score
// at least that much) and drops the pairs whose s + e < t.  At the end the
// survivors are re-scored with the exact chain and sorted by (score desc,
// id asc).  A query whose bound admits more pairs than its buffer holds
// (scores crowded around its k-th) is marked uncertain instead, and an
// exact pass computes its row: the chain score of every base row, kept in
// a workgroup's candidate list (the key (score, id) is unique, so the list
// is cut to exactly k and never overflows).
//
// Layout: a workgroup of 4 waves takes 64 queries, 16 per wave, and
// streams all base rows 64 at a time; the waves are independent (each
// owns its queries' buffers and thresholds) and read the same base rows,
// which the CU's cache shares.  Query fragments stay in registers (A
// operand: lane l holds query row l % 16, elements 32c + 8(l / 16) .. +7);
// base fragments (B operand) are loaded per 32-element chunk; the result
// lane l holds queries 4(l / 16) .. +3 against base row l % 16 of the
// 16-row tile.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "pg_device.hpp"
#include "pg_kernels.h"

namespace pg {

constexpr uint32_t kTopkCap = kTopkBufferPairs;  // candidate slots per query
constexpr uint32_t kTopkStep = 64;        // base rows per step
constexpr uint32_t kTopkQueries = 64;     // queries per workgroup
constexpr uint32_t kTopkUncertain = 0xFFFFFFFFu;  // count of an uncertain row
constexpr uint32_t kExactCap = 2048;      // exact pass: candidate slots

// Per-row bound terms: the query side (h_q, r_q, g_c n_q) and the base
// side (r_x + g_a h_x, h_x + r_x, n_x) of the pair bound e above.
struct TopkSide {
  float a, b, c;
};
__device__ __forceinline__ TopkSide topk_query_side(float n, float r,
                                                   float gam_c) {
  return TopkSide{n + r, r, gam_c * n};
}
__device__ __forceinline__ TopkSide topk_base_side(float n, float r,
                                                  float gam_a) {
  const float h = n + r;
  return TopkSide{r + gam_a * h, h + r, n};
}
__device__ __forceinline__ float topk_err(const TopkSide& q,
                                          const TopkSide& x) {
  return 1.01f * (q.a * x.a + q.b * x.b + q.c * x.c);
}

// Sorts the n (<= kTopkCap) buffered pairs of one query by lower bound,
// descending, sets the threshold to the k-th lower bound (when there are k
// pairs) and keeps the pairs whose upper bound reaches it.  One wave; its
// own LDS scratch (keys, pairs).
__device__ void topk_compact(uint2* __restrict__ qbuf, uint32_t n,
                             uint32_t k, TopkSide qs,
                             const float* __restrict__ bnorm,
                             const float* __restrict__ bres, float gam_a,
                             uint64_t* __restrict__ skey,
                             uint2* __restrict__ spair, float* theta,
                             uint32_t* count) {
  const uint32_t lane = lane_id();
  if (n > kTopkCap) n = kTopkCap;  // an overflow, already flagged
  auto err = [&](uint32_t id) {
    return topk_err(qs, topk_base_side(bnorm[id], bres[id], gam_a));
  };
  for (uint32_t j = lane; j < kTopkCap; j += kWave) {
    uint64_t key = kPadKey;
    if (j < n) {
      const uint2 e = qbuf[j];
      spair[j] = e;
      const float lo = __uint_as_float(e.y) - err(e.x);
      key = (uint64_t(ordered_bits(-lo)) << 32) | j;
    }
    skey[j] = key;
  }
  wave_sync();
  bitonic_sort<uint64_t, true>(skey, kTopkCap);
  float t = -FLT_MAX;
  if (n >= k) {
    const uint2 e = spair[uint32_t(skey[k - 1])];
    t = __uint_as_float(e.y) - err(e.x);
  }
  if (t < *theta) t = *theta;  // a lower bound found earlier still holds
  // keep the pairs whose upper bound reaches t, in lower-bound order
  uint32_t m = 0;
  for (uint32_t j0 = 0; j0 < n; j0 += kWave) {
    const uint32_t j = j0 + lane;
    bool keep = false;
    uint2 e = make_uint2(0u, 0u);
    if (j < n) {
      e = spair[uint32_t(skey[j])];
      keep = __uint_as_float(e.y) + err(e.x) >= t;
    }
    const uint64_t km = __ballot(keep);
    if (keep) qbuf[m + __popcll(km & lanemask_lt())] = e;
    m += __popcll(km);
  }
  wave_sync();
  if (lane == 0) {
    *theta = t;
    *count = m;
  }
  wave_sync();
}

template <int NCH>
__global__ __launch_bounds__(256) void topk_screen_kernel(
    const __bf16* __restrict__ q16, const float* __restrict__ qnorm,
    const float* __restrict__ qres, uint32_t nq,
    const __bf16* __restrict__ b16, const float* __restrict__ bnorm,
    const float* __restrict__ bres, uint32_t nb, uint32_t dpad, float gam_a,
    float gam_c, uint32_t k, uint2* __restrict__ buf,
    uint32_t* __restrict__ count_out) {
  __shared__ float theta_s[kTopkQueries];
  __shared__ uint32_t count_s[kTopkQueries];
  __shared__ uint32_t unc_s[kTopkQueries];
  __shared__ uint64_t skey_s[4][kTopkCap];
  __shared__ uint2 spair_s[4][kTopkCap];
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t m = lane & 15u, kq = lane >> 4;
  const uint32_t nch = dpad / 32u;
  const uint32_t qw = blockIdx.x * kTopkQueries + 16u * wave;  // first query
  float* theta = theta_s + 16u * wave;
  uint32_t* count = count_s + 16u * wave;
  uint32_t* unc = unc_s + 16u * wave;
  if (lane < 16u) {
    theta[lane] = -FLT_MAX;
    count[lane] = 0u;
    unc[lane] = 0u;
  }
  // query fragments (A operand) and the bound factors of this lane's rows
  bf16x8 a[NCH];
  {
    const uint32_t qa = qw + m;
    const bf16x8* row = reinterpret_cast<const bf16x8*>(
        q16 + size_t(qa < nq ? qa : 0u) * dpad + 8u * kq);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (qa < nq && uint32_t(c) < nch) {
        a[c] = row[4 * c];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[c][j] = (__bf16)0.0f;
      }
    }
  }
  TopkSide eq[4];
  bool qok[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t qi = qw + 4u * kq + uint32_t(r);
    qok[r] = qi < nq;
    eq[r] = qok[r] ? topk_query_side(qnorm[qi], qres[qi], gam_c)
                   : TopkSide{0.0f, 0.0f, 0.0f};
  }
  wave_sync();
  for (uint32_t t0 = 0; t0 < nb; t0 += kTopkStep) {
    f32x4 acc[4];
    TopkSide xs[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      const uint32_t col = t0 + 16u * uint32_t(j) + m;
      xs[j] = col < nb ? topk_base_side(bnorm[col], bres[col], gam_a)
                       : TopkSide{0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (uint32_t(c) < nch) {
        bf16x8 b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t col = t0 + 16u * uint32_t(j) + m;
          if (col < nb) {
            b[j] = *reinterpret_cast<const bf16x8*>(
                b16 + size_t(col) * dpad + 32u * uint32_t(c) + 8u * kq);
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) b[j][e] = (__bf16)0.0f;
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b[j], acc[j],
                                                           0, 0, 0);
      }
    }
    // keep the pairs whose upper bound reaches their query's threshold
    float th[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) th[r] = theta[4u * kq + uint32_t(r)];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t col = t0 + 16u * uint32_t(j) + m;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float s = acc[j][r];
        if (col < nb && qok[r] && s + topk_err(eq[r], xs[j]) >= th[r]) {
          const uint32_t ql = 4u * kq + uint32_t(r);
          const uint32_t slot = atomicAdd(&count[ql], 1u);
          // the buffer holds the pairs within the bound of a query's k-th
          // score; past kTopkCap of them the row goes to the exact pass
          if (slot < kTopkCap)
            buf[size_t(qw + ql) * kTopkCap + slot] =
                make_uint2(col, __float_as_uint(s));
          else
            unc[ql] = 1u;
        }
      }
    }
    wave_sync();
    // an uncertain row takes no more pairs (its threshold closes); a query
    // gains at most kTopkStep pairs per step: compact the buffers that
    // could overflow in the next one
    if (lane < 16u && unc[lane]) {
      theta[lane] = FLT_MAX;
      count[lane] = 0u;
    }
    wave_sync();
    const bool full =
        lane < 16u && count[lane] > kTopkCap - kTopkStep;
    uint64_t fm = __ballot(full);
    while (fm) {
      const uint32_t ql = uint32_t(__ffsll((long long)fm) - 1);
      fm &= fm - 1;
      const uint32_t qi = qw + ql;
      topk_compact(buf + size_t(qi) * kTopkCap, count[ql], k,
                   topk_query_side(qnorm[qi], qres[qi], gam_c), bnorm, bres,
                   gam_a, skey_s[wave], spair_s[wave], &theta[ql],
                   &count[ql]);
    }
  }
  // final compaction of every query of this wave
  for (uint32_t ql = 0; ql < 16u; ++ql) {
    const uint32_t qi = qw + ql;
    if (qi >= nq) break;
    if (unc[ql]) {
      if (lane == 0) count_out[qi] = kTopkUncertain;
      continue;
    }
    topk_compact(buf + size_t(qi) * kTopkCap, count[ql], k,
                 topk_query_side(qnorm[qi], qres[qi], gam_c), bnorm, bres,
                 gam_a, skey_s[wave], spair_s[wave], &theta[ql], &count[ql]);
    if (lane == 0) count_out[qi] = count[ql];
  }
}

// The uncertain rows, listed for the exact pass.
__global__ void topk_uncertain_kernel(const uint32_t* __restrict__ count,
                                      uint32_t nq, uint32_t* __restrict__ list,
                                      uint32_t* __restrict__ n_list) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq;
       q += gridDim.x * blockDim.x)
    if (count[q] == kTopkUncertain) list[atomicAdd(n_list, 1u)] = q;
}

// Ascending bitonic sort of n (a power of two) keys in LDS by a whole
// workgroup.
__device__ void block_bitonic_sort(uint64_t* a, uint32_t n) {
  const uint32_t half = n >> 1;
  for (uint32_t kk = 2; kk <= n; kk <<= 1) {
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
        const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const uint32_t p = i | j;
        const uint64_t x = a[i], y = a[p];
        const bool asc = (i & kk) == 0;
        if ((x > y) == asc) {
          a[i] = y;
          a[p] = x;
        }
      }
      __syncthreads();
    }
  }
}

// Exact pass for one uncertain row per workgroup (list[blockIdx.x]): the
// chain score of every base row, key (score desc, id asc); candidates that
// beat the current k-th key enter the list, which is sorted and cut to k
// whenever the next step could fill it.
__global__ __launch_bounds__(256) void topk_exact_rows_kernel(
    const float* __restrict__ queries, const float* __restrict__ base,
    uint32_t nb, uint32_t dim, uint32_t k, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ n_list, int32_t* __restrict__ out) {
  __shared__ uint64_t key[kExactCap];
  __shared__ uint32_t cnt;
  __shared__ uint64_t thr;
  if (blockIdx.x >= *n_list) return;
  const uint32_t q = list[blockIdx.x];
  const float* xq = queries + size_t(q) * dim;
  if (threadIdx.x == 0) {
    cnt = 0u;
    thr = kPadKey;
  }
  __syncthreads();
  auto cut = [&](uint32_t keep) {
    for (uint32_t j = cnt + threadIdx.x; j < kExactCap; j += blockDim.x)
      key[j] = kPadKey;
    __syncthreads();
    block_bitonic_sort(key, kExactCap);
    if (threadIdx.x == 0) {
      if (cnt > keep) cnt = keep;
      if (cnt >= k) thr = key[k - 1];
    }
    __syncthreads();
  };
  for (uint32_t r0 = 0; r0 < nb; r0 += blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    if (r < nb) {
      const float d = chain_dist(xq, base + size_t(r) * dim, dim);
      const uint64_t kk = (uint64_t(ordered_bits(d)) << 32) | r;
      if (kk < thr) key[atomicAdd(&cnt, 1u)] = kk;
    }
    __syncthreads();
    if (cnt > kExactCap - blockDim.x) cut(k);
  }
  cut(k);
  for (uint32_t j = threadIdx.x; j < k; j += blockDim.x)
    out[size_t(q) * k + j] = j < cnt ? int32_t(uint32_t(key[j])) : -1;
}

// Re-scores the survivors of each query with the exact chain and writes the
// k best ids (score desc, id asc), -1 past the survivors.  One wave per
// query.
__global__ __launch_bounds__(64) void topk_final_kernel(
    const float* __restrict__ queries, const float* __restrict__ base,
    uint32_t dim, const uint2* __restrict__ buf,
    const uint32_t* __restrict__ count, uint32_t nq, uint32_t k,
    int32_t* __restrict__ out) {
  __shared__ uint64_t key[kTopkCap < 64 ? 64 : kTopkCap];
  const uint32_t q = blockIdx.x;
  const uint32_t lane = lane_id();
  if (q >= nq) return;
  const uint32_t n = count[q];
  if (n == kTopkUncertain) return;  // written by the exact pass
  const float* xq = queries + size_t(q) * dim;
  const uint2* qb = buf + size_t(q) * kTopkCap;
  uint32_t np2 = 64;
  while (np2 < n) np2 <<= 1;
  for (uint32_t j = lane; j < np2; j += kWave) {
    uint64_t kk = kPadKey;
    if (j < n) {
      const uint32_t id = qb[j].x;
      // chain_dist is -chain(q, x); the largest score is the smallest key
      const float d = chain_dist(xq, base + size_t(id) * dim, dim);
      kk = (uint64_t(ordered_bits(d)) << 32) | id;
    }
    key[j] = kk;
  }
  __syncthreads();
  bitonic_sort(key, np2);
  for (uint32_t j = lane; j < k; j += kWave)
    out[size_t(q) * k + j] = j < n ? int32_t(uint32_t(key[j])) : -1;
}

hipError_t launch_exact_topk(const float* queries, uint32_t nq,
                             const float* base, uint32_t nb, uint32_t dim,
                             uint32_t k, const void* q16, const float* qnorm,
                             const float* qres, const void* b16,
                             const float* bnorm, const float* bres,
                             uint32_t dpad, float gam_a, float gam_c,
                             uint2* buf, uint32_t* count, uint32_t* list,
                             uint32_t* n_list, int32_t* out, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  const uint32_t blocks = (nq + kTopkQueries - 1) / kTopkQueries;
  const uint32_t nch = dpad / 32u;
  if (nch <= 8) {
    topk_screen_kernel<8><<<blocks, 256, 0, s>>>(
        reinterpret_cast<const __bf16*>(q16), qnorm, qres, nq,
        reinterpret_cast<const __bf16*>(b16), bnorm, bres, nb, dpad, gam_a,
        gam_c, k, buf, count);
  } else if (nch <= 16) {
    topk_screen_kernel<16><<<blocks, 256, 0, s>>>(
        reinterpret_cast<const __bf16*>(q16), qnorm, qres, nq,
        reinterpret_cast<const __bf16*>(b16), bnorm, bres, nb, dpad, gam_a,
        gam_c, k, buf, count);
  } else {
    return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  topk_final_kernel<<<nq, 64, 0, s>>>(queries, base, dim, buf, count, nq, k,
                                      out);
  e = launch_fill_u32(n_list, 1, 0u, s);
  if (e != hipSuccess) return e;
  topk_uncertain_kernel<<<(nq + 255) / 256 < 4096 ? (nq + 255) / 256 : 4096,
                          256, 0, s>>>(count, nq, list, n_list);
  // one workgroup per uncertain row; the others return at once
  topk_exact_rows_kernel<<<nq, 256, 0, s>>>(queries, base, nb, dim, k, list,
                                            n_list, out);
  return hipGetLastError();
}

}  // namespace pg
