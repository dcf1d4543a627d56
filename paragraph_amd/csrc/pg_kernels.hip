// Index-build kernels for MI355X (gfx950, wave64), plain FP32 path.
//
// One wavefront (= one 64-lane workgroup) per node or per query.  Distances
// are computed with explicit FMAs in exactly the reference's summation
// orders, so neighbour ids match the CPU oracle bit for bit:
//   prune / merge: sequential chain over the row  (Gbuilder.cuh:785-797)
//   search:        32 lane-strided chains + 16/8/4/2/1 tree
//                                     (Gsearcher.cuh:958-976, 1023-1035)
// Ordering rules where the reference depends on scheduling (C1-C3 in the
// oracle's header) are implemented deterministically here.
#include "pg_device.hpp"
#include "pg_kernels.h"
#include "pg_search.hpp"

namespace pg {

// ---------------------------------------------------------------- utilities

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n,
                                uint32_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, uint64_t n,
                                unsigned long long v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Sums the per-slot counters into one row (called between pipeline stages,
// so that per-stage deltas can be reported).
__global__ __launch_bounds__(256) void snapshot_stats_kernel(
    const unsigned long long* __restrict__ stats,
    unsigned long long* __restrict__ row) {
  __shared__ unsigned long long part[256];
  const uint32_t t = threadIdx.x;
  const uint32_t f = t % kStatFields, s0 = t / kStatFields;  // 16 x 16
  unsigned long long v = 0;
  for (uint32_t s = s0; s < kStatSlots; s += 256 / kStatFields)
    v += stats[s * kStatFields + f];
  part[t] = v;
  __syncthreads();
  if (t < kStatFields) {
    unsigned long long acc = 0;
    for (uint32_t k = 0; k < 256 / kStatFields; ++k) acc += part[k * kStatFields + t];
    row[t] = acc;
  }
}

// ------------------------------------------------------------ K0: top-1 match
// match_top1_kernel_v0 (Gbuilder.cuh:104-135).  The reference lets an
// atomicCAS pick the owner of a contested base row; here the lowest query id
// wins (rule C1) via atomicMin on a per-base winner array.  The row itself
// is never materialised: the prune step reads gt[winner][1..gt_k-2].
// gt holds the rows of queries q0 .. q0 + n_query - 1 (a chunk of the gt
// when the host streams it in).
__global__ void match_winner_kernel(const uint32_t* __restrict__ gt,
                                    uint32_t q0, uint32_t n_query,
                                    uint32_t gt_k, uint32_t n_base,
                                    uint32_t* __restrict__ winner) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n_query;
       q += gridDim.x * blockDim.x) {
    uint32_t b = gt[size_t(q) * gt_k];
    if (b < n_base) atomicMin(&winner[b], q0 + q);
  }
}

// ------------------------------------------------- RNG occlusion (K2 and K3)
// Strict RNG over a sorted list held in LDS (Gbuilder.cuh:1150-1198 and
// :1500-1553).  `skey[i]` = (ordered distance << 32 | original position),
// `sid[pos]` = id.  kept[0..nk) already holds the first kept entry.  Each
// 64-candidate block is resolved with lane = candidate: every lane first
// tests its candidate against the entries kept before the block, then the
// block is walked in order, each newly kept entry testing the later lanes.
// Every candidate therefore compares against kept entries in insertion
// order and stops at the first occluder, as the reference does; the pair
// counter only counts comparisons the reference would make.
template <bool kStopIds>
__device__ uint32_t rng_select(const float* __restrict__ base, uint32_t n_base,
                               uint32_t dim, const uint64_t* skey,
                               const uint32_t* sid, uint32_t n_scan,
                               uint32_t cap, uint32_t* kept, uint32_t nk,
                               unsigned long long* pair_total) {
  const uint32_t lane = lane_id();
  unsigned long long pairs = 0;
  bool stop = false;
  for (uint32_t blk = 1; blk < n_scan && !stop && nk < cap; blk += kWave) {
    const uint32_t pos = blk + lane;
    const bool in = pos < n_scan;
    uint32_t my_id = kTomb;
    float my_d = FLT_MAX;
    if (in) {
      uint64_t k = skey[pos];
      my_id = sid[uint32_t(k)];
      my_d = from_ordered(uint32_t(k >> 32));
    }
    uint32_t lim = kWave;
    if (kStopIds) {
      uint64_t sm = __ballot(in && is_stop_id(my_id, n_base));
      if (sm) {
        lim = __ffsll((long long)sm) - 1;
        stop = true;
      }
    }
    const bool active = in && lane < lim;
    bool occl = false;
    uint32_t my_pairs = 0;
    const float* own = base + size_t(active ? my_id : 0) * dim;
    for (uint32_t k = 0; k < nk; ++k) {
      const uint32_t kid = kept[k];
      if (active && !occl) {
        float pd = chain_dist(base + size_t(kid) * dim, own, dim);
        ++my_pairs;
        occl = pd < my_d;
      }
    }
    int last = -1;
    int cut = kWave;
    while (true) {
      uint64_t m = __ballot(active && !occl && int(lane) > last);
      if (m == 0) break;
      const int j = __ffsll((long long)m) - 1;
      const uint32_t kid = __shfl(my_id, j, 64);
      if (lane == 0) kept[nk] = kid;
      ++nk;
      last = j;
      if (nk == cap) {
        cut = j;
        break;
      }
      if (active && !occl && int(lane) > j) {
        float pd = chain_dist(base + size_t(kid) * dim, own, dim);
        ++my_pairs;
        occl = pd < my_d;
      }
    }
    pairs += wave_sum<unsigned long long>(
        (active && int(lane) <= cut) ? my_pairs : 0u);
    wave_sync();
  }
  *pair_total += pairs;
  return nk;
}

// chains give the block against each 16-wide tile of kept entries and
// against the block itself (BF16 rows, or the hi + lo split of each row for a
// ~30x tighter bound; FP32 accumulation).  A screened value s decides a
// comparison only when the bound on |s - chain| (pair_bound: rounding of the
// rows, FP32 accumulation, the chain's own error) puts both sides of d_i out
// of reach;
// otherwise the exact FP32 chain is computed (counted as a fallback).  So
// every decision equals the plain FP32 path's.
// 1: occluded for sure, 0: not occluded for sure, -1: undecided.
__device__ __forceinline__ int screen_cmp(float s, float d, float e) {
  if (s + e < d) return 1;
  if (s - e >= d) return 0;
  return -1;
}

// Bound on |screened - exact chain| for the pair (x, y), with n = |x| and
// r = |x - bf16(x)| per row.  Plain screen: x.y - x^.y^ = x^.r_y + r_x.y^ +
// r_x.r_y, at most h_x r_y + r_x h_y + r_x r_y with h = n + r >= |x^|
// (Cauchy-Schwarz); plus the MFMA accumulation gam_a h_x h_y (products of
// BF16 values are exact in FP32) and the chain's own gam_c n_x n_y; 1% margin
// for the FP32 norms and this evaluation.  Split screen: eps_c n_x n_y.
__device__ __forceinline__ float pair_bound(const ScreenArgs& sc, float nx,
                                            float rx, float ny, float ry) {
  if (sc.rnorms == nullptr) return sc.eps_c * nx * ny;
  const float hx = nx + rx, hy = ny + ry;
  return 1.01f * (sc.gam_a * hx * hy + hx * ry + rx * hy + rx * ry +
                  sc.gam_c * nx * ny);
}

// Bound on |x^.y^ + x^.bf16(y - y^) - chain(x, y)|: the row x in BF16
// against a column y carried with its BF16 residual (two MFMA columns, summed
// in FP32).  Error terms: r_x.y <= r_x n_y; x^.(r_y - bf16(r_y)) <= h_x (2^-8
// r_y + sqrt(dpad) 2^-126) (BF16 rounding of the residual, flushed
// subnormals included); the FP32 accumulation of both chains and their sum,
// (gam_a + 2^-24) h_x (h_y + 1.01 r_y); the exact chain's own gam_c n_x n_y;
// 1% margin for the FP32 norms and this evaluation.
__device__ __forceinline__ float pair_bound_res(const ScreenArgs& sc, float nx,
                                                float rx, float ny, float ry) {
  const float hx = nx + rx, hy = ny + ry;
  return 1.01f * (rx * ny + hx * (0.00390625f * ry + 1e-36f) +
                  (sc.gam_a + 5.9604645e-8f) * hx * (hy + 1.01f * ry) +
                  sc.gam_c * nx * ny);
}

// cand_nr: when given, the (norm, residual norm) of candidate j = the low
// word of its key, in LDS (else they are read from sc per block).
template <bool kStopIds, int NCH, int NKT, bool kSplit>
__device__ uint32_t rng_select_screened(
    const float* __restrict__ base, const ScreenArgs& sc, uint32_t n_base,
    uint32_t dim, const uint64_t* skey, const uint32_t* sid, uint32_t n_scan,
    uint32_t cap, uint32_t* kept, float2* kept_nr, uint32_t nk, float* gs,
    unsigned long long* pair_total, unsigned long long* fb_total,
    const float2* cand_nr = nullptr) {
  constexpr uint32_t GW = NKT * 16 + 16;  // gs row width
  const uint32_t lane = lane_id();
  const uint32_t m = lane & 15u, kq = lane >> 4;
  const uint32_t nch = sc.dpad / 32;
  unsigned long long pairs = 0, fb = 0;
  bool stop = false;
  for (uint32_t blk = 1; blk < n_scan && !stop && nk < cap; blk += 16) {
    const uint32_t pos = blk + m;
    const bool in = pos < n_scan;
    uint32_t my_id = kTomb, my_j = 0;
    float my_d = FLT_MAX;
    if (in) {
      const uint64_t k = skey[pos];
      my_j = uint32_t(k);
      my_id = sid[my_j];
      my_d = from_ordered(uint32_t(k >> 32));
    }
    uint32_t lim = 16;
    if (kStopIds) {
      const uint64_t sm = __ballot(lane < 16 && in && is_stop_id(my_id, n_base));
      if (sm) {
        lim = __ffsll((long long)sm) - 1;
        stop = true;
      }
    }
    const bool active = in && m < lim;
    const uint32_t nk0 = nk;
    // ---- Gram tiles
    // Plain: BF16 rows.  Split (kSplit): rows x = hi + lo as two BF16 copies
    // and x.y taken as hi.hi + lo.hi + hi.lo (three MFMAs).  The chunks are
    // taken in register groups of at most 8: the block's fragments of one
    // group stay in registers while every kept tile accumulates.
    constexpr int CG = NCH < 8 ? NCH : 8;
    f32x4 acc_t[NKT], acc_s = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < NKT; ++t) acc_t[t] = acc_s;
    const size_t own_off = size_t(active ? my_id : 0u) * sc.dpad + 8u * kq;
#pragma unroll
    for (int g0 = 0; g0 < NCH; g0 += CG) {
      bf16x8 ah[CG], al[kSplit ? CG : 1];
      {
        const bf16x8* rh = reinterpret_cast<const bf16x8*>(sc.base16 + own_off);
        const bf16x8* rl = reinterpret_cast<const bf16x8*>(
            kSplit ? sc.base16lo + own_off : sc.base16 + own_off);
#pragma unroll
        for (int c = 0; c < CG; ++c) {
          const bool ld = active && uint32_t(g0 + c) < nch;
          if (ld) {
            ah[c] = rh[4 * (g0 + c)];
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) ah[c][j] = (__bf16)0.0f;
          }
          if constexpr (kSplit) {
            if (ld) {
              al[c] = rl[4 * (g0 + c)];
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) al[c][j] = (__bf16)0.0f;
            }
          }
        }
      }
#pragma unroll
      for (int t = 0; t < NKT; ++t) {
        if (16u * t < nk0) {
          const uint32_t kidx = 16u * t + m;
          const bool kv = kidx < nk0;
          const size_t off = size_t(kv ? kept[kidx] : 0u) * sc.dpad + 8u * kq;
          const bf16x8* bh = reinterpret_cast<const bf16x8*>(sc.base16 + off);
          const bf16x8* bl = reinterpret_cast<const bf16x8*>(
              kSplit ? sc.base16lo + off : sc.base16 + off);
#pragma unroll
          for (int c = 0; c < CG; ++c) {
            const bool ld = kv && uint32_t(g0 + c) < nch;
            bf16x8 h;
            if (ld) {
              h = bh[4 * (g0 + c)];
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) h[j] = (__bf16)0.0f;
            }
            acc_t[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], h,
                                                               acc_t[t], 0, 0,
                                                               0);
            if constexpr (kSplit) {
              bf16x8 lw;
              if (ld) {
                lw = bl[4 * (g0 + c)];
              } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) lw[j] = (__bf16)0.0f;
              }
              acc_t[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  al[c], h, acc_t[t], 0, 0, 0);
              acc_t[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  ah[c], lw, acc_t[t], 0, 0, 0);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CG; ++c) {
        acc_s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], ah[c], acc_s, 0,
                                                        0, 0);
        if constexpr (kSplit) {
          acc_s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c], ah[c], acc_s,
                                                          0, 0, 0);
          acc_s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], al[c], acc_s,
                                                          0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NKT; ++t)
      if (16u * t < nk0)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          gs[(4u * kq + j) * GW + 16u * t + m] = acc_t[t][j];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      gs[(4u * kq + j) * GW + NKT * 16u + m] = acc_s[j];
    wave_sync();
    // ---- resolution: lane i < 16 owns block candidate i (the first
    // occluder in kept order ends a candidate's scan early)
    const bool own = lane < 16 && active;
    float my_norm = 0.0f, my_rn = 0.0f;
    if (own) {
      if (cand_nr) {
        my_norm = cand_nr[my_j].x;
        my_rn = cand_nr[my_j].y;
      } else {
        my_norm = sc.norms[my_id];
        my_rn = sc.rnorms ? sc.rnorms[my_id] : 0.0f;
      }
    }
    const float* own_row = base + size_t(own ? my_id : 0u) * dim;
    bool occl = false;
    uint32_t my_pairs = 0, my_fb = 0;
    if (own) {
      for (uint32_t k = 0; k < nk0; ++k) {
        ++my_pairs;
        const float2 knr = kept_nr[k];
        int dec = screen_cmp(-gs[m * GW + k], my_d,
                             pair_bound(sc, my_norm, my_rn, knr.x, knr.y));
        if (dec < 0) {
          ++my_fb;
          dec = chain_dist(base + size_t(kept[k]) * dim, own_row, dim) < my_d;
        }
        if (dec) {
          occl = true;
          break;
        }
      }
    }
    int last = -1, cut = 16;
    while (true) {
      const uint64_t msk = __ballot(own && !occl && int(lane) > last);
      if (msk == 0) break;
      const int j = __ffsll((long long)msk) - 1;
      const uint32_t kid = __shfl(my_id, j, 64);
      const float kn = __shfl(my_norm, j, 64);
      const float kr = __shfl(my_rn, j, 64);
      if (lane == 0) {
        kept[nk] = kid;
        kept_nr[nk] = make_float2(kn, kr);
      }
      ++nk;
      last = j;
      if (nk == cap) {
        cut = j;
        break;
      }
      if (own && !occl && int(lane) > j) {
        ++my_pairs;
        int dec = screen_cmp(-gs[m * GW + NKT * 16u + j], my_d,
                             pair_bound(sc, my_norm, my_rn, kn, kr));
        if (dec < 0) {
          ++my_fb;
          dec = chain_dist(base + size_t(kid) * dim, own_row, dim) < my_d;
        }
        if (dec) occl = true;
      }
    }
    pairs += wave_sum<unsigned long long>((own && int(lane) <= cut) ? my_pairs : 0u);
    fb += wave_sum<unsigned long long>(own ? my_fb : 0u);
    wave_sync();
  }
  *pair_total += pairs;
  *fb_total += fb;
  return nk;
}

// --------------------------------------------------------- K2: prune (pruned)
// fusion_prune_reverse_kernel_v0 (Gbuilder.cuh:1105-1244), one wave per node.
// Candidates: rows[b][0..D) or, in top-1 mode (rows == nullptr), the row the
// match step would have written: gt[winner[b]][1..gt_k-2] then two tombs
// (Gbuilder.cuh:121-132).  tomb / self / INT_MAX / out-of-range get FLT_MAX;
// the list is sorted by (distance, position) which equals the reference's
// stable bubble sort; the node is skipped if the first entry is a stop id;
// then strict RNG up to mp kept.  The reverse edges are built afterwards
// from the pruned lists (reverse_* kernels).
// NCH > 0 selects the screened RNG (BF16 MFMA tiles, NCH 32-wide chunks,
// NKT 16-wide kept tiles); NCH == 0 is the plain FP32 path.
template <uint32_t DP2, int NCH, int NKT, bool kSplit>
__global__ __launch_bounds__(64) void prune_kernel(
    const float* __restrict__ base, uint32_t n_base, uint32_t dim,
    const uint32_t* __restrict__ rows, uint32_t D,
    const uint32_t* __restrict__ gt, const uint32_t* __restrict__ winner,
    uint32_t gt_k, uint32_t mp, uint32_t* __restrict__ pruned,
    uint32_t* __restrict__ pcnt, unsigned long long* __restrict__ stats,
    ScreenArgs sc, const uint32_t* __restrict__ order) {
  __shared__ uint64_t skey[DP2];
  __shared__ uint32_t sid[DP2];
  __shared__ uint32_t kept[128];
  __shared__ float2 kept_nr[NCH ? 128 : 1];
  __shared__ float gs[NCH ? 16 * (NKT * 16 + 16) : 1];
  // nodes in the locality order when one is given (rows are per node, so
  // the launch order changes no result)
  const uint32_t b = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t lane = lane_id();
  __shared__ uint16_t vpos[DP2];
  const uint32_t w = rows ? 0u : winner[b];
  const float* xb = base + size_t(b) * dim;
  // ids; invalid slots get FLT_MAX; valid positions are compacted so that
  // every lane of the distance pass has a real candidate
  uint32_t nv = 0;
  for (uint32_t j0 = 0; j0 < DP2; j0 += kWave) {
    const uint32_t j = j0 + lane;
    uint32_t id = kTomb;
    if (j < D) {
      if (rows) {
        id = rows[size_t(b) * D + j];
      } else if (w != kTomb && j + 2 < gt_k) {
        id = gt[size_t(w) * gt_k + j + 1];
      }
    }
    sid[j] = id;
    const bool valid =
        j < D && !(id == kTomb || id == b || id == kMaskedTomb || id >= n_base);
    skey[j] = j < D ? ((uint64_t(ordered_bits(FLT_MAX)) << 32) | j) : kPadKey;
    const uint64_t vm = __ballot(valid);
    if (valid) vpos[nv + __popcll(vm & lanemask_lt())] = uint16_t(j);
    nv += __popcll(vm);
  }
  wave_sync();
  for (uint32_t v = lane; v < nv; v += kWave) {
    const uint32_t j = vpos[v];
    const float d = chain_dist(xb, base + size_t(sid[j]) * dim, dim);
    skey[j] = (uint64_t(ordered_bits(d)) << 32) | j;
  }
  wave_sync();
  sort_list<true>(skey, DP2);
  unsigned long long pairs = 0, fbk = 0;
  uint32_t nk = 0;
  const uint32_t id0 = sid[uint32_t(skey[0])];
  if (!is_stop_id(id0, n_base)) {
    if (lane == 0) {
      kept[0] = id0;
      if (NCH)
        kept_nr[0] = make_float2(sc.norms[id0],
                                 sc.rnorms ? sc.rnorms[id0] : 0.0f);
    }
    wave_sync();
    if constexpr (NCH == 0)
      nk = rng_select<true>(base, n_base, dim, skey, sid, D, mp, kept, 1,
                            &pairs);
    else
      nk = rng_select_screened<true, NCH, NKT, kSplit>(base, sc, n_base, dim, skey,
                                               sid, D, mp, kept, kept_nr, 1,
                                               gs, &pairs, &fbk);
  }
  for (uint32_t j = lane; j < nk; j += kWave) pruned[size_t(b) * mp + j] = kept[j];
  if (lane == 0) pcnt[b] = nk;
  if (lane == 0) {
    stat_add(stats, ST_PRUNE_DIST, nv);
    stat_add(stats, ST_PRUNE_PAIRS, pairs);
    stat_add(stats, ST_SCREEN_FALLBACK, fbk);
  }
}

// ------------------------------------------ K2 on link rows: kept-major sweep
// The strict RNG keeps a candidate when no earlier kept entry occludes it.
// Taken kept-major, that reads: the remaining entry with the smallest key is
// always kept (every entry that an earlier kept entry occludes has already
// been removed), and it removes every remaining entry that it occludes.
// That gives prune_kernel's kept lists exactly (same entries in the same
// order, the same cap, the same stop at the first tomb / INT_MAX /
// out-of-range id in key order, self entries at FLT_MAX by position), while
// each sweep reads every remaining candidate row once.  On inner-product
// rows the first kept entry occludes almost every candidate, so one sweep
// replaces prune_kernel's distance pass and RNG pass.
//
// kScreen (BF16 rows): one MFMA per chunk and 16 rows gives the screened
// product of each row with the current kept entry (B columns 0 and 2: the
// entry and its BF16 residual) and, on the first sweep, with the node itself
// (columns 1 and 3: the screened d).  pair_bound_res makes both intervals; an
// open decision falls back to exact FP32 chains, and only entries that may
// hold the smallest key get an exact d.  The first sweep is made against the row's first candidate
// (search rows are sorted, so that is nearly always the smallest key) and is
// repeated against the right entry when it was not.
// !kScreen: fused lane-per-row chains give the exact d and the exact pair
// distance in one pass over each row; the pair counter then equals
// prune_kernel's candidate-major count.

// chain_dist(x, p) and chain_dist(y, p) in one pass over p: two independent
// FMA chains in index order, bit-equal to the two separate calls.
__device__ __forceinline__ void chain_dist2(const float* __restrict__ x,
                                            const float* __restrict__ y,
                                            const float* __restrict__ p,
                                            uint32_t dim, float& dx,
                                            float& dy) {
  float s = 0.0f, t = 0.0f;
  if ((dim & 3u) == 0) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* y4 = reinterpret_cast<const float4*>(y);
    const float4* p4 = reinterpret_cast<const float4*>(p);
    const uint32_t n4 = dim >> 2;
    uint32_t i = 0;
    for (; i + 4 <= n4; i += 4) {
      float4 a[4], c[4], v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v[u] = p4[i + u];
        a[u] = x4[i + u];
        c[u] = y4[i + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s = __builtin_fmaf(a[u].x, v[u].x, s);
        t = __builtin_fmaf(c[u].x, v[u].x, t);
        s = __builtin_fmaf(a[u].y, v[u].y, s);
        t = __builtin_fmaf(c[u].y, v[u].y, t);
        s = __builtin_fmaf(a[u].z, v[u].z, s);
        t = __builtin_fmaf(c[u].z, v[u].z, t);
        s = __builtin_fmaf(a[u].w, v[u].w, s);
        t = __builtin_fmaf(c[u].w, v[u].w, t);
      }
    }
    for (; i < n4; ++i) {
      const float4 a = x4[i], c = y4[i], v = p4[i];
      s = __builtin_fmaf(a.x, v.x, s);
      t = __builtin_fmaf(c.x, v.x, t);
      s = __builtin_fmaf(a.y, v.y, s);
      t = __builtin_fmaf(c.y, v.y, t);
      s = __builtin_fmaf(a.z, v.z, s);
      t = __builtin_fmaf(c.z, v.z, t);
      s = __builtin_fmaf(a.w, v.w, s);
      t = __builtin_fmaf(c.w, v.w, t);
    }
  } else {
    for (uint32_t i = 0; i < dim; ++i) {
      const float v = p[i];
      s = __builtin_fmaf(x[i], v, s);
      t = __builtin_fmaf(y[i], v, t);
    }
  }
  dx = -s;
  dy = -t;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = shfl_xor_u64(v, uint32_t(o));
    v = w < v ? w : v;
  }
  return v;
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, xor_lane_dyn(v, uint32_t(o)));
  return v;
}

// float4s per lane in flight in the exact quad chains (a quad holds 4x as
// many): the sweep's 16 take rows of up to 256 floats in one load round;
// the merge's 13 take up to 208 (dim 200), as 16 would push the dim-512
// split-screen merge past 256 VGPRs.
constexpr int kSweepChainBatch = 16;
constexpr int kMergeChainBatch = 13;

// Elements 8 lane .. +7 of row id (lanes < 4 NCH; zero past dim and past
// the row's dpad / 32 chunks), as the screened sweep's B columns take them.
template <int NCH>
__device__ __forceinline__ void sweep_row_octet(const float* __restrict__ base,
                                                uint32_t dim,
                                                const ScreenArgs& sc,
                                                uint32_t id, float (&v)[8]) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = 0.0f;
  if (lane < 4u * NCH)
    load_query_octet(base + size_t(id) * dim, dim, lane < 4u * (sc.dpad / 32u),
                     8u * lane, v);
}

// LDS of one sweep (one wave).  Entries e < ne; key of e = (ordered d,
// position << 9 | e).
template <uint32_t DP2, int NCH, bool kScreen>
struct SweepLds {
  // B columns: 0 kept entry, 1 node, 2 / 3 their BF16 residuals, stored
  // in the order (0, 2, 1, 3): slots 0-1 hold the kept entry's two columns,
  // slots 2-3 the node's.  Once the first sweep has used the node's
  // columns, slots 2-3 hold its FP32 row, and once a sweep has used the
  // kept entry's, slots 0-1 may hold that entry's FP32 row (2 slots = dpad
  // floats).
  bf16x8 frag[kScreen ? 4 * NCH * 4 : 1];
  float2 enr[kScreen ? DP2 : 1];  // (norm, residual norm)
  uint32_t eid[DP2];
  float dv[DP2];   // exact d when (ef & 1), else the screened d
  float sv[DP2];   // pair value against the current kept entry
  uint32_t kept[128];
  uint16_t epos[DP2];
  uint16_t alive[DP2];
  // remover's kept index; 0x8000 | t for kept entry t (the exact kernel's
  // pair count)
  uint16_t rk[kScreen ? 1 : DP2];
  uint16_t wl[kScreen ? DP2 : 1];  // entries waiting for an exact chain
  uint8_t ef[DP2];   // bit 0: d exact; bit 1: kept or removed
};

// Kept-major strict RNG over the entries set up in L (eid, epos, alive = 0 ..
// ne-1, ef = 1 and dv = exact d for entries whose d is known, rk = 0xFFFF,
// enr for the screened kernel), ne > 0.  c is the speculative first kept
// entry.  The screened kernel takes the node's norms (nb_in, rb_in) and the
// octets of the node's and c's FP32 rows that sweep_row_octet loaded (the
// caller issues those loads early).  Entries whose key is not below kstop
// end the scan.  Returns the kept count (at most cap), kept ids in L.kept.
// pairs: exact kernel, the candidate-major comparison count; screened
// kernel, the screened pair decisions.  fb: exact chains taken.
template <uint32_t DP2, int NCH, bool kScreen>
__device__ __forceinline__ uint32_t sweep_select(
    SweepLds<DP2, NCH, kScreen>& L, const float* __restrict__ base,
    uint32_t dim, uint32_t node, const ScreenArgs& sc, uint32_t ne,
    uint32_t cap, uint64_t kstop, uint32_t c, float nb_in, float rb_in,
    const float (&node_oct)[8], const float (&c_oct)[8],
    unsigned long long& pairs_out, unsigned long long& fb_out) {
  constexpr int G = NCH <= 8 ? 2 : 1;            // 16-row groups in flight
  constexpr uint32_t NF = kScreen ? NCH * 4 : 1;  // fragments per row
  const uint32_t lane = lane_id();
  const uint32_t m = lane & 15u, kq = lane >> 4;
  const float* xb = base + size_t(node) * dim;
  const uint32_t nch = kScreen ? sc.dpad / 32u : 0u;
  float nb = 0.0f, rb = 0.0f;
  unsigned long long my_pairs = 0, my_fb = 0;
  // screened kernel, column col (0 kept entry / 1 node) from the FP32 row:
  // the BF16 column and its residual column (col + 2): hi = bf16(v) (the
  // conversion that made the BF16 copy), lo = bf16(v - hi)
  // The node's FP32 row, elements 8 lane .. +7 of lanes < NF (zero past
  // dim), kept in registers from its fragment load until the first sweep
  // has used the node's slots.
  float nrow[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  // column col from the octet v of lane < NF
  auto store_frag = [&](uint32_t col, const float (&v)[8]) {
    if constexpr (kScreen) {
      if (lane < NF) {
        bf16x8 h, l;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          h[j] = (__bf16)v[j];
          l[j] = (__bf16)(v[j] - (float)h[j]);
          if (col == 1u) nrow[j] = v[j];
        }
        L.frag[2u * col * NF + lane] = h;
        L.frag[(2u * col + 1u) * NF + lane] = l;
      }
    }
  };
  auto load_frag = [&](uint32_t col, uint32_t id) {
    if constexpr (kScreen) {
      float v[8];
      sweep_row_octet<NCH>(base, dim, sc, id, v);
      store_frag(col, v);
    }
  };
  // FP32 rows in LDS for the exact chains (screened kernel): the node's in
  // slots 2-3, the current kept entry's in slots 0-1
  float* const xnode = reinterpret_cast<float*>(L.frag + 2u * NF);
  float* const xkept = reinterpret_cast<float*>(L.frag);
  auto stage_row = [&](float* dst, uint32_t id) {
    const float* x = base + size_t(id) * dim;
    if ((dim & 3u) == 0) {
      for (uint32_t f = lane; f < (dim >> 2); f += kWave)
        reinterpret_cast<float4*>(dst)[f] =
            reinterpret_cast<const float4*>(x)[f];
    } else {
      for (uint32_t i = lane; i < dim; i += kWave) dst[i] = x[i];
    }
  };
  // Exact chains of the n entries listed in L.wl, 16 at a time (a quad per
  // entry): d against the node row (slots 2-3; sets ef bit 0) or, with
  // xs, the pair value against the kept row (slots 0-1) into L.sv.  Every
  // lane calls it (quad_chain needs the whole wave).
  auto wl_chains = [&](uint32_t n, bool xs) {
    for (uint32_t b0 = 0; b0 < n; b0 += kWave / 4u) {
      const uint32_t idx = b0 + (lane >> 2);
      const bool on = idx < n;
      const uint32_t e = on ? L.wl[idx] : 0u;
      const float v = quad_chain<kSweepChainBatch>(
          xs ? xkept : xnode, base + size_t(L.eid[e]) * dim, on, dim);
      if (on && (lane & 3u) == 0) {
        ++my_fb;
        if (xs) {
          L.sv[e] = v;
        } else {
          L.dv[e] = v;
          L.ef[e] |= 1u;
        }
      }
    }
    wave_sync();
  };
  // exact d of entry e (one lane)
  auto exact_d = [&](uint32_t e) -> float {
    return chain_dist(xb, base + size_t(L.eid[e]) * dim, dim);
  };
  // bound of the screened d of entry e
  auto eb_of = [&](uint32_t e) -> float {
    return pair_bound_res(sc, L.enr[e].x, L.enr[e].y, nb, rb);
  };
  if constexpr (kScreen) {
    nb = nb_in;
    rb = rb_in;
  }
  store_frag(1, node_oct);
  auto key_of = [&](uint32_t e) -> uint64_t {
    return (uint64_t(ordered_bits(L.dv[e])) << 32) |
           (uint32_t(L.epos[e]) << 9) | e;
  };
  auto lo_of = [&](uint32_t e) -> float {
    if constexpr (kScreen)
      if (!(L.ef[e] & 1u)) return L.dv[e] - eb_of(e);
    return L.dv[e];
  };
  auto hi_of = [&](uint32_t e) -> float {
    if constexpr (kScreen)
      if (!(L.ef[e] & 1u)) return L.dv[e] + eb_of(e);
    return L.dv[e];
  };
  uint32_t nal = ne;

  // Pair values of every alive entry against entry c (the screened kernel
  // reads c's fragments from L.frag); on the first sweep also d.
  auto sweep = [&](uint32_t c, bool first) {
    if constexpr (kScreen) {
      for (uint32_t g0 = 0; g0 < nal; g0 += 16u * G) {
        bf16x8 a[G][NCH];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const uint32_t idx = g0 + 16u * g + m;
          const bool ok = idx < nal;
          const uint32_t id = ok ? L.eid[L.alive[idx]] : 0u;
          const bf16x8* row = reinterpret_cast<const bf16x8*>(
              sc.base16 + size_t(id) * sc.dpad + 8u * kq);
          // lanes past nal read row 0 (a valid row): their output rows are
          // never stored, so only the chunk bound (uniform) decides a load
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            if (uint32_t(ch) < nch) {
              a[g][ch] = row[4 * ch];
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) a[g][ch][j] = (__bf16)0.0f;
            }
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
          // MFMA column m (< 4) is storage slot ((m & 1) << 1) | (m >> 1)
          const uint32_t slot = m < 4u ? (((m & 1u) << 1) | (m >> 1)) : 0u;
          const bf16x8* bf = L.frag + slot * NF + kq;
          __asm__ volatile("" ::: "memory");
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g][ch], bf[4 * ch],
                                                          acc, 0, 0, 0);
          // lane l holds rows 4(l/16) .. +3 of column l%16; columns 0 / 1
          // take their residual columns 2 / 3 from lane l + 2
          float res[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) res[j] = __shfl(acc[j], int(lane) + 2, 64);
          if (m < 2u) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const uint32_t idx = g0 + 16u * g + 4u * kq + j;
              if (idx < nal) {
                const uint32_t e = L.alive[idx];
                const float v = -(acc[j] + res[j]);
                if (m == 0u)
                  L.sv[e] = v;
                else if (first && !(L.ef[e] & 1u))
                  L.dv[e] = v;
              }
            }
          }
        }
      }
    } else {
      const float* xc = base + size_t(L.eid[c]) * dim;
      for (uint32_t i = lane; i < nal; i += kWave) {
        const uint32_t e = L.alive[i];
        const float* xp = base + size_t(L.eid[e]) * dim;
        float s;
        if (first && !(L.ef[e] & 1u)) {
          float d;
          chain_dist2(xb, xc, xp, dim, d, s);
          L.dv[e] = d;
          L.ef[e] |= 1u;
        } else {
          s = chain_dist(xc, xp, dim);
        }
        L.sv[e] = s;
      }
    }
    wave_sync();
  };

  // The alive entry with the smallest exact key.  Only entries whose lower
  // bound does not exceed the smallest upper bound can hold it; with one
  // such entry no exact d is needed, otherwise those get exact chains.
  auto pick = [&]() -> uint32_t {
    float mh = INFINITY;
    for (uint32_t i = lane; i < nal; i += kWave)
      mh = fminf(mh, hi_of(L.alive[i]));
    mh = wave_min_f32(mh);
    uint32_t nt = 0, first_t = 0;
    for (uint32_t i0 = 0; i0 < nal; i0 += kWave) {
      const uint32_t i = i0 + lane;
      uint32_t e = 0;
      bool in = false;
      if (i < nal) {
        e = L.alive[i];
        in = !(lo_of(e) > mh);
      }
      const uint64_t tm = __ballot(in);
      if (tm && nt == 0) first_t = __shfl(e, __ffsll((long long)tm) - 1, 64);
      nt += __popcll(tm);
    }
    if (nt == 1) return first_t;
    if constexpr (kScreen) {
      // exact d for the screened entries of that set, as quad chains
      uint32_t nw = 0;
      for (uint32_t i0 = 0; i0 < nal; i0 += kWave) {
        const uint32_t i = i0 + lane;
        uint32_t e = 0;
        bool need = false;
        if (i < nal) {
          e = L.alive[i];
          need = !(lo_of(e) > mh) && !(L.ef[e] & 1u);
        }
        const uint64_t bm = __ballot(need);
        if (need) L.wl[nw + __popcll(bm & lanemask_lt())] = uint16_t(e);
        nw += __popcll(bm);
      }
      wave_sync();
      if (nw > 0) wl_chains(nw, false);
    }
    uint64_t best = kPadKey;
    for (uint32_t i = lane; i < nal; i += kWave) {
      const uint32_t e = L.alive[i];
      if (lo_of(e) > mh) continue;
      if (!(L.ef[e] & 1u)) {
        L.dv[e] = exact_d(e);
        L.ef[e] |= 1u;
        ++my_fb;
      }
      const uint64_t k = key_of(e);
      best = k < best ? k : best;
    }
    best = wave_min_u64(best);
    wave_sync();
    return uint32_t(best & 511u);
  };

  uint32_t nk = 0, clast = 0;
  store_frag(0, c_oct);
  wave_sync();
  sweep(c, true);
  if constexpr (kScreen) {
    // the node's columns are used: its FP32 row takes their slots
    if (lane < NF) {
      reinterpret_cast<float4*>(xnode)[2u * lane] =
          make_float4(nrow[0], nrow[1], nrow[2], nrow[3]);
      reinterpret_cast<float4*>(xnode)[2u * lane + 1u] =
          make_float4(nrow[4], nrow[5], nrow[6], nrow[7]);
    }
    wave_sync();
  }
  const uint32_t c1 = pick();
  if (c1 != c) {
    c = c1;
    load_frag(0, L.eid[c]);
    wave_sync();
    sweep(c, false);
  }
  while (true) {
    // stop rule: an entry whose key is not below kstop ends the scan
    if (kstop != kPadKey && !(L.ef[c] & 1u) && !(hi_of(c) < FLT_MAX)) {
      if constexpr (kScreen) {
        if (lane == 0) L.wl[0] = uint16_t(c);
        wave_sync();
        wl_chains(1, false);
      } else {
        if (lane == 0) {
          L.dv[c] = exact_d(c);
          L.ef[c] |= 1u;
          ++my_fb;
        }
        wave_sync();
      }
    }
    if ((L.ef[c] & 1u) && !(key_of(c) < kstop)) break;
    if (lane == 0) {
      L.kept[nk] = L.eid[c];
      if constexpr (!kScreen) L.rk[c] = uint16_t(0x8000u | nk);
      L.ef[c] |= 2u;
    }
    ++nk;
    clast = c;
    wave_sync();
    if (nk == cap) break;
    // every alive entry against the new kept entry c
    float nc = 0.0f, rc = 0.0f;
    if constexpr (kScreen) {
      nc = L.enr[c].x;
      rc = L.enr[c].y;
    }
    if constexpr (kScreen) {
      // Screened decisions (ef bit 1: removed).  An open interval needs the
      // exact d first, then, if screen_cmp stays open, the exact pair value
      // (ef bit 2 marks those); both run as quad chains over lists in L.wl.
      auto decide = [&](uint32_t e, float s, float ec, bool& need_d) {
        if (s + ec < lo_of(e)) {
          L.ef[e] |= 2u;
        } else if (s - ec >= hi_of(e)) {
        } else if (!(L.ef[e] & 1u)) {
          need_d = true;
        } else {
          const int dec = screen_cmp(s, L.dv[e], ec);
          if (dec < 0) L.ef[e] |= 4u;
          else if (dec) L.ef[e] |= 2u;
        }
      };
      uint32_t nw = 0;
      for (uint32_t i0 = 0; i0 < nal; i0 += kWave) {
        const uint32_t i = i0 + lane;
        bool need = false;
        uint32_t e = 0;
        if (i < nal) {
          e = L.alive[i];
          if (e != c) {
            ++my_pairs;
            decide(e, L.sv[e],
                   pair_bound_res(sc, L.enr[e].x, L.enr[e].y, nc, rc), need);
          }
        }
        const uint64_t bm = __ballot(need);
        if (need) L.wl[nw + __popcll(bm & lanemask_lt())] = uint16_t(e);
        nw += __popcll(bm);
      }
      wave_sync();
      if (nw > 0) {
        wl_chains(nw, false);
        for (uint32_t j = lane; j < nw; j += kWave) {
          const uint32_t e = L.wl[j];
          bool unused = false;
          decide(e, L.sv[e],
                 pair_bound_res(sc, L.enr[e].x, L.enr[e].y, nc, rc), unused);
        }
        wave_sync();
      }
      uint32_t ns = 0;
      for (uint32_t i0 = 0; i0 < nal; i0 += kWave) {
        const uint32_t i = i0 + lane;
        const uint32_t e = i < nal ? L.alive[i] : 0u;
        const bool pend = i < nal && (L.ef[e] & 4u);
        const uint64_t bm = __ballot(pend);
        if (pend) L.wl[ns + __popcll(bm & lanemask_lt())] = uint16_t(e);
        ns += __popcll(bm);
      }
      wave_sync();
      if (ns > 0) {
        // the kept entry's columns are used: its FP32 row takes their slots
        stage_row(xkept, L.eid[c]);
        wave_sync();
        wl_chains(ns, true);
        for (uint32_t j = lane; j < ns; j += kWave) {
          const uint32_t e = L.wl[j];
          L.ef[e] = uint8_t((L.ef[e] & ~4u) | (L.sv[e] < L.dv[e] ? 2u : 0u));
        }
      }
      wave_sync();
    } else {
      for (uint32_t i = lane; i < nal; i += kWave) {
        const uint32_t e = L.alive[i];
        if (e == c) continue;
        if (L.sv[e] < L.dv[e]) {
          L.ef[e] |= 2u;
          L.rk[e] = uint16_t(nk - 1);
        }
      }
      wave_sync();
    }
    uint32_t n2 = 0;
    for (uint32_t i0 = 0; i0 < nal; i0 += kWave) {
      const uint32_t i = i0 + lane;
      const uint32_t e = i < nal ? L.alive[i] : 0u;
      const bool keep = i < nal && !(L.ef[e] & 2u);
      const uint64_t km = __ballot(keep);
      if (keep) L.alive[n2 + __popcll(km & lanemask_lt())] = uint16_t(e);
      n2 += __popcll(km);
    }
    nal = n2;
    wave_sync();
    if (nal == 0) break;
    c = pick();
    load_frag(0, L.eid[c]);
    wave_sync();
    sweep(c, false);
  }
  if constexpr (!kScreen) {
    // candidate-major pair count: entries up to the last processed key
    const uint64_t kend = nk == cap ? key_of(clast) : kstop;
    for (uint32_t e = lane; e < ne; e += kWave) {
      const uint32_t r = L.rk[e];
      if (r == 0xFFFFu) continue;
      if (r & 0x8000u)
        my_pairs += r & 0x7FFFu;
      else if (key_of(e) < kend)
        my_pairs += r + 1u;
    }
  }
  pairs_out = wave_sum<unsigned long long>(my_pairs);
  fb_out = wave_sum<unsigned long long>(my_fb);
  return nk;
}

template <uint32_t DP2, int NCH, bool kScreen>
__global__ __launch_bounds__(64) void prune_sweep_kernel(
    const float* __restrict__ base, uint32_t n_base, uint32_t dim,
    const uint32_t* __restrict__ rows, uint32_t D, uint32_t mp,
    uint32_t* __restrict__ pruned, uint32_t* __restrict__ pcnt,
    unsigned long long* __restrict__ stats, ScreenArgs sc,
    const uint32_t* __restrict__ order) {
  __shared__ SweepLds<DP2, NCH, kScreen> L;
  // nodes in the locality order when one is given (rows are per node, so
  // the launch order changes no result)
  const uint32_t b = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t lane = lane_id();
  // The loads are issued in dependency levels, each level's loads in
  // flight together: the node's norms and row octet with the row ids, then
  // the entries' norms with the first kept entry's row octet.
  float nb = 0.0f, rb = 0.0f;
  float noct[8], coct[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) noct[j] = coct[j] = 0.0f;
  if constexpr (kScreen) {
    nb = sc.norms[b];
    rb = sc.rnorms[b];
    sweep_row_octet<NCH>(base, dim, sc, b, noct);
  }
  constexpr int NJ = DP2 / kWave;  // row slices of 64 ids
  uint32_t idr[NJ];
#pragma unroll
  for (int r = 0; r < NJ; ++r) {
    const uint32_t j = 64u * r + lane;
    idr[r] = (64u * r < D && j < D) ? rows[size_t(b) * D + j] : kTomb;
  }
  // Entries: valid candidates, and self entries before the first stop id
  // (prune_kernel sorts them at FLT_MAX by position).
  uint32_t jstop = D;
#pragma unroll
  for (int r = NJ - 1; r >= 0; --r) {
    const uint32_t j = 64u * r + lane;
    const uint64_t sm = __ballot(j < D && is_stop_id(idr[r], n_base));
    if (sm) jstop = 64u * r + __ffsll((long long)sm) - 1;
  }
  // the entries' norms, and the speculative first kept entry: the first
  // valid candidate in row order (else entry 0, a self entry)
  float2 nrr[kScreen ? NJ : 1];
  uint32_t cid = b, jc = D;
#pragma unroll
  for (int r = 0; r < NJ; ++r) {
    const uint32_t j = 64u * r + lane;
    const uint32_t id = idr[r];
    const bool stop = j >= D || is_stop_id(id, n_base);
    if constexpr (kScreen)
      nrr[r] = stop ? make_float2(0.0f, 0.0f)
                    : make_float2(sc.norms[id], sc.rnorms[id]);
    const uint64_t vm = __ballot(!stop && id != b);
    if (vm && jc == D) {
      const int l = __ffsll((long long)vm) - 1;
      jc = 64u * r + l;
      cid = __builtin_amdgcn_readlane(id, l);
    }
  }
  if constexpr (kScreen) sweep_row_octet<NCH>(base, dim, sc, cid, coct);
  uint32_t ne = 0, nv = 0, c = 0xFFFFFFFFu;
#pragma unroll
  for (int r = 0; r < NJ; ++r) {
    const uint32_t j = 64u * r + lane;
    const uint32_t id = idr[r];
    const bool stop = j >= D || is_stop_id(id, n_base);
    const bool self = !stop && id == b;
    const bool valid = !stop && !self;
    const bool inc = valid || (self && j < jstop);
    const uint64_t im = __ballot(inc);
    nv += __popcll(__ballot(valid));
    if (jc >= 64u * r && jc < 64u * r + 64u)
      c = ne + __popcll(im & ((1ull << (jc - 64u * r)) - 1ull));
    if (inc) {
      const uint32_t e = ne + __popcll(im & lanemask_lt());
      L.eid[e] = id;
      L.epos[e] = uint16_t(j);
      L.alive[e] = uint16_t(e);
      L.ef[e] = self ? 1u : 0u;
      L.dv[e] = self ? FLT_MAX : 0.0f;
      L.rk[e] = 0xFFFFu;
      if constexpr (kScreen) L.enr[e] = nrr[r];
    }
    ne += __popcll(im);
  }
  wave_sync();
  const uint64_t kstop =
      jstop < D ? ((uint64_t(ordered_bits(FLT_MAX)) << 32) | (jstop << 9))
                : kPadKey;
  uint32_t nk = 0;
  unsigned long long pairs = 0, fbk = 0;
  if (ne > 0) {
    if (c == 0xFFFFFFFFu) c = 0;
    nk = sweep_select<DP2, NCH, kScreen>(L, base, dim, b, sc, ne, mp, kstop,
                                         c, nb, rb, noct, coct, pairs, fbk);
  }
  for (uint32_t j = lane; j < nk; j += kWave) pruned[size_t(b) * mp + j] = L.kept[j];
  if (lane == 0) {
    pcnt[b] = nk;
    stat_add(stats, ST_PRUNE_DIST, nv);
    stat_add(stats, ST_PRUNE_PAIRS, pairs);
    stat_add(stats, ST_SCREEN_FALLBACK, fbk);
  }
}

// ------------------------------------------------------ reverse edges (CSR)
// add_reverse (Gbuilder.cuh:1020-1038) keeps whichever sources win the
// atomicAdd race; here every edge b -> c is recorded in a CSR keyed by c and
// the merge step keeps the cap smallest sources (rule C2).
//
// Inner-product graphs have hubs (one node can be kept by 2% of all nodes),
// and device-scope atomics on one address serialize across the XCDs.  So
// each block takes a contiguous range of kRevNodes nodes, aggregates their
// edges in an LDS hash table (target -> count) and issues one global atomic
// per distinct target.  An edge whose target finds no slot within kRevProbe
// probes goes straight to global memory; because slots are never freed, all
// edges of one target either share one slot or all overflow, in both passes.
constexpr uint32_t kRevSlots = 8192;
constexpr uint32_t kRevNodes = 1024;
constexpr uint32_t kRevProbe = 32;
constexpr uint32_t kRevEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t rev_hash(uint32_t c) {
  return (c * 2654435761u) >> 19;  // 13 bits
}

// Slot of target c (inserted if absent), or kRevSlots when the probe run is
// full.
__device__ __forceinline__ uint32_t rev_insert(uint32_t* keys, uint32_t c) {
  uint32_t h = rev_hash(c);
  for (uint32_t p = 0; p < kRevProbe; ++p) {
    const uint32_t prev = atomicCAS(&keys[h], kRevEmpty, c);
    if (prev == kRevEmpty || prev == c) return h;
    h = (h + 1u) & (kRevSlots - 1u);
  }
  return kRevSlots;
}

__device__ __forceinline__ uint32_t rev_find(const uint32_t* keys,
                                             uint32_t c) {
  uint32_t h = rev_hash(c);
  for (uint32_t p = 0; p < kRevProbe; ++p) {
    const uint32_t k = keys[h];
    if (k == c) return h;
    if (k == kRevEmpty) return kRevSlots;
    h = (h + 1u) & (kRevSlots - 1u);
  }
  return kRevSlots;
}

__global__ __launch_bounds__(256) void reverse_count_kernel(
    const uint32_t* __restrict__ pruned, const uint32_t* __restrict__ pcnt,
    uint32_t n_base, uint32_t mp, uint32_t* __restrict__ deg) {
  __shared__ uint32_t keys[kRevSlots];
  __shared__ uint32_t cnt[kRevSlots];
  for (uint32_t i = threadIdx.x; i < kRevSlots; i += blockDim.x) {
    keys[i] = kRevEmpty;
    cnt[i] = 0u;
  }
  __syncthreads();
  const uint32_t b0 = blockIdx.x * kRevNodes;
  const uint32_t b1 = min(b0 + kRevNodes, n_base);
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    const uint32_t n = pcnt[b];
    const uint32_t* row = pruned + size_t(b) * mp;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t c = row[j];
      if (c >= n_base) continue;
      const uint32_t h = rev_insert(keys, c);
      if (h < kRevSlots)
        atomicAdd(&cnt[h], 1u);
      else
        atomicAdd(&deg[c], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kRevSlots; i += blockDim.x)
    if (keys[i] != kRevEmpty) atomicAdd(&deg[keys[i]], cnt[i]);
}

// Two passes over the block's edges: the first builds the same table (and
// places overflow edges directly), then one global atomic per slot reserves
// the target's run, and the second pass ranks each edge inside its run.
__global__ __launch_bounds__(256) void reverse_fill_kernel(
    const uint32_t* __restrict__ pruned, const uint32_t* __restrict__ pcnt,
    uint32_t n_base, uint32_t mp, const uint32_t* __restrict__ off,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ srcs) {
  __shared__ uint32_t keys[kRevSlots];
  __shared__ uint32_t cnt[kRevSlots];
  __shared__ uint32_t run[kRevSlots];
  for (uint32_t i = threadIdx.x; i < kRevSlots; i += blockDim.x) {
    keys[i] = kRevEmpty;
    cnt[i] = 0u;
  }
  __syncthreads();
  const uint32_t b0 = blockIdx.x * kRevNodes;
  const uint32_t b1 = min(b0 + kRevNodes, n_base);
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    const uint32_t n = pcnt[b];
    const uint32_t* row = pruned + size_t(b) * mp;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t c = row[j];
      if (c >= n_base) continue;
      const uint32_t h = rev_insert(keys, c);
      if (h < kRevSlots) {
        atomicAdd(&cnt[h], 1u);
      } else {
        const uint32_t slot = atomicAdd(&cursor[c], 1u);
        srcs[off[c] + slot] = b;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kRevSlots; i += blockDim.x) {
    const uint32_t c = keys[i];
    if (c != kRevEmpty) run[i] = off[c] + atomicAdd(&cursor[c], cnt[i]);
    cnt[i] = 0u;
  }
  __syncthreads();
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    const uint32_t n = pcnt[b];
    const uint32_t* row = pruned + size_t(b) * mp;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t c = row[j];
      if (c >= n_base) continue;
      const uint32_t h = rev_find(keys, c);
      if (h < kRevSlots) srcs[run[h] + atomicAdd(&cnt[h], 1u)] = b;
    }
  }
}

// Exclusive scan, 3 passes: per-block scan of 1024 entries, scan of the
// block sums (one block), add-back.  in has n entries, out n+1.
constexpr uint32_t kScanBlock = 1024;

__global__ __launch_bounds__(256) void scan_blocks_kernel(
    const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
    uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t wsum[4];
  const uint32_t base_i = blockIdx.x * kScanBlock;
  const uint32_t t = threadIdx.x;
  uint32_t v[4];
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = base_i + t * 4 + k;
    v[k] = i < n ? in[i] : 0u;
    run += v[k];
  }
  // inclusive scan of per-thread totals across the block
  uint32_t x = run;
  const uint32_t lane = t & 63u, wv = t >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= uint32_t(o)) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  uint32_t wprefix = 0;
  for (uint32_t k = 0; k < wv; ++k) wprefix += wsum[k];
  uint32_t excl = wprefix + x - run;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = base_i + t * 4 + k;
    if (i < n) out[i] = excl;
    excl += v[k];
  }
  if (t == 255) block_sums[blockIdx.x] = wprefix + x;
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(
    uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ total) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (uint32_t base_i = 0; base_i < nb; base_i += 1024) {
    const uint32_t i = base_i + t;
    const uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o, 64);
      if (lane >= uint32_t(o)) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    uint32_t wprefix = 0;
    for (uint32_t k = 0; k < wv; ++k) wprefix += wsum[k];
    const uint32_t c = carry;
    if (i < nb) sums[i] = c + wprefix + x - v;
    __syncthreads();
    if (t == 1023) carry = c + wprefix + x;
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

__global__ void scan_add_kernel(uint32_t* __restrict__ out, uint32_t n,
                                const uint32_t* __restrict__ block_sums) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += block_sums[i / kScanBlock];
}

// ------------------------------------------------------------ K3: merge
// merge_and_sort + fusion_merge_sort_prune_kernel (Gbuilder.cuh:1433-1601),
// one wave per node c.  Candidates: the cap smallest reverse sources of c
// in ascending id (C2), then c's pruned list.  Stable distance sort, list[0]
// kept unconditionally, strict RNG to `deg`, non-strict fill when !strict
// (Gbuilder.cuh:1559-1594), tomb fill.
// Sources are selected in LDS chunks of kSegChunk ids (longer segments take
// several sort passes): 256 keeps the screened kernel at 8.3 KB of LDS, four
// waves per SIMD instead of three (BF16 merge 4.8 / 5.1 -> 4.2 / 4.5 ms per
// round on 1M x 200).
constexpr uint32_t kSegChunk = 256;
// Lists of at most this many sources / candidates are ordered by rank
// placement (wave_rank) instead of a 64-wide bitonic sort: on the 1M x 200
// value path the merge's lists average under 4 candidates.
constexpr uint32_t kRankSortMax = 32;

template <int NCH, int NKT, bool kSplit>
__global__ __launch_bounds__(64) void merge_kernel(
    const float* __restrict__ base, uint32_t n_base, uint32_t dim,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ srcs,
    uint32_t cap, const uint32_t* __restrict__ pruned,
    const uint32_t* __restrict__ pcnt, uint32_t mp, uint32_t deg,
    int strict, uint32_t* __restrict__ out,
    unsigned long long* __restrict__ stats, ScreenArgs sc,
    const uint32_t* __restrict__ order) {
  // 16-byte aligned: it also stages the node's FP32 row as float4s
  __shared__ __align__(16) uint32_t buf[2 * kSegChunk];
  __shared__ uint64_t skey[128];
  __shared__ uint32_t sid[128];
  __shared__ uint32_t kept[128];
  __shared__ float2 kept_nr[NCH ? 128 : 1];
  __shared__ float2 cand_nr[NCH ? 128 : 1];  // (norm, residual) by index
  __shared__ float gs[NCH ? 16 * (NKT * 16 + 16) : 1];
  __shared__ uint8_t added[128];
  // nodes in the locality order when one is given (one output row per
  // node, so the launch order changes no result)
  const uint32_t c = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t lane = lane_id();
  // The loads are issued in dependency levels, each level's loads in
  // flight together: the segment bounds, the pruned count and the node's
  // row (for the quad chains below), then the first 64 sources and pruned
  // ids.
  const uint32_t s0 = off[c], s1 = off[c + 1];
  const uint32_t p = pcnt[c];
  const float* xc = base + size_t(c) * dim;
  const bool xrow = dim <= 2 * kSegChunk && (dim & 3u) == 0;
  float4 xr[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t f = lane + 64u * k;
    xr[k] = (xrow && f < (dim >> 2))
                ? reinterpret_cast<const float4*>(xc)[f]
                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const uint32_t seg = s1 - s0;
  const uint32_t r = seg < cap ? seg : cap;
  const uint32_t pv = lane < p ? pruned[size_t(c) * mp + lane] : kTomb;
  // ---- select the r smallest sources, ascending
  if (seg <= kRankSortMax) {
    // a short segment (most nodes): each source placed at its rank
    const uint32_t v = lane < seg ? srcs[s0 + lane] : kTomb;
    const uint32_t rk = wave_rank(v, seg);
    if (lane < seg) buf[rk] = v;
    wave_sync();
  } else if (seg <= kSegChunk) {
    uint32_t n2 = 64;
    while (n2 < seg) n2 <<= 1;
    for (uint32_t j = lane; j < n2; j += kWave) buf[j] = j < seg ? srcs[s0 + j] : kTomb;
    wave_sync();
    if (seg > 1) sort_list<true>(buf, n2);
  } else if (cap > 0) {
    // A long segment (a hub: up to millions of sources).  buf[0, cap) holds
    // the cap smallest ids seen so far, ascending, and thr its largest: a
    // later id can only enter if it is below thr.  Such ids are staged
    // after the list and merged in (one sort) when the next chunk could
    // overflow the stage, so each chunk costs a compare per id.  Source ids
    // are distinct (b -> c is one edge per pruned row), and with repeats
    // the strict test still keeps the cap smallest of the multiset.
    for (uint32_t j = lane; j < kSegChunk; j += kWave) buf[j] = srcs[s0 + j];
    wave_sync();
    bitonic_sort<uint32_t, true>(buf, kSegChunk);
    uint32_t thr = buf[cap - 1];
    uint32_t ns = 0;  // staged ids at buf[cap, cap + ns)
    auto fold = [&]() {
      for (uint32_t j = cap + ns + lane; j < 2 * kSegChunk; j += kWave)
        buf[j] = kTomb;
      wave_sync();
      bitonic_sort<uint32_t, true>(buf, 2 * kSegChunk);
      thr = buf[cap - 1];
      ns = 0;
    };
    for (uint32_t done = kSegChunk; done < seg; done += kSegChunk) {
      uint32_t v[kSegChunk / kWave];
#pragma unroll
      for (uint32_t k = 0; k < kSegChunk / kWave; ++k) {
        const uint32_t g = done + k * kWave + lane;
        v[k] = g < seg ? srcs[s0 + g] : kTomb;
      }
#pragma unroll
      for (uint32_t k = 0; k < kSegChunk / kWave; ++k) {
        const bool in = v[k] < thr;
        const uint64_t im = __ballot(in);
        if (in) buf[cap + ns + __popcll(im & lanemask_lt())] = v[k];
        ns += __popcll(im);
      }
      if (cap + ns + kSegChunk > 2 * kSegChunk) fold();
    }
    if (ns > 0) fold();
    // buf[cap, kSegChunk) may hold staged or sorted ids past the list; the
    // candidate list below reads only buf[0, r)
  }
  // ---- candidate list: reverse part then pruned part
  const uint32_t n = r + p;
  if (lane < p) sid[r + lane] = pv;
  for (uint32_t i = 64u + lane; i < p; i += kWave)
    sid[r + i] = pruned[size_t(c) * mp + i];
  for (uint32_t j = lane; j < 128; j += kWave) {
    if (j < r) {
      sid[j] = buf[j];
    } else if (j >= n) {
      sid[j] = kTomb;
      skey[j] = kPadKey;
    }
  }
  wave_sync();
  // the screened RNG's norms of every candidate, loaded with the rows of the
  // exact d below
  if constexpr (NCH != 0) {
    for (uint32_t j = lane; j < n; j += kWave) {
      const uint32_t id = sid[j];
      cand_nr[j] = make_float2(sc.norms[id], sc.rnorms ? sc.rnorms[id] : 0.0f);
    }
  }
  // exact d of every candidate: up to 32 candidates as quad chains (four
  // lanes per candidate, its whole row in flight) against the node's row,
  // staged in buf (dead once the sources are in sid); longer lists (where
  // quads would take more passes than one lane per candidate takes load
  // rounds) and wider rows take one lane per candidate
  if (n <= 32 && xrow) {
    float* xs = reinterpret_cast<float*>(buf);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t f = lane + 64u * k;
      if (f < (dim >> 2)) reinterpret_cast<float4*>(xs)[f] = xr[k];
    }
    wave_sync();
    for (uint32_t b0 = 0; b0 < n; b0 += kWave / 4u) {
      const uint32_t j = b0 + (lane >> 2);
      const bool on = j < n;
      const float d = quad_chain<kMergeChainBatch>(
          xs, base + size_t(on ? sid[j] : 0u) * dim, on, dim);
      if (on && (lane & 3u) == 0)
        skey[j] = (uint64_t(ordered_bits(d)) << 32) | j;
    }
  } else {
    // a fixed trip count (128 / 64 lanes) keeps both chains of a lane in
    // flight together
#pragma unroll
    for (uint32_t j = lane; j < 128; j += kWave)
      if (j < n)
        skey[j] = (uint64_t(ordered_bits(chain_dist(
                      xc, base + size_t(sid[j]) * dim, dim))) << 32) | j;
  }
  wave_sync();
  uint32_t* o = out + size_t(c) * deg;
  if (n == 0) {
    for (uint32_t j = lane; j < deg; j += kWave) o[j] = kTomb;
    return;
  }
  if (n <= kRankSortMax) {
    // a short list (most nodes): each key placed at its rank (keys are
    // distinct: the low word is the candidate's position)
    const uint64_t k = lane < n ? skey[lane] : kPadKey;
    const uint32_t rk = wave_rank_distinct(k, n);
    wave_sync();
    if (lane < n) skey[rk] = k;
    wave_sync();
  } else {
    // keys past n are padding, so the sort covers the next power of two
    // >= n (at least 64: one register slice per lane)
    uint32_t nk2 = 64;
    while (nk2 < n) nk2 <<= 1;
    sort_list<true>(skey, nk2);
  }
  unsigned long long pairs = 0, fbk = 0;
  if (lane == 0) {
    kept[0] = sid[uint32_t(skey[0])];
    if (NCH) kept_nr[0] = cand_nr[uint32_t(skey[0])];
  }
  wave_sync();
  uint32_t nk;
  if constexpr (NCH == 0)
    nk = rng_select<false>(base, n_base, dim, skey, sid, n, deg, kept, 1,
                           &pairs);
  else
    nk = rng_select_screened<false, NCH, NKT, kSplit>(
        base, sc, n_base, dim, skey, sid, n, deg, kept, kept_nr, 1, gs,
        &pairs, &fbk, cand_nr);
  if (!strict && nk < deg) {
    // mark the first list slot equal to each kept id, then append the
    // unmarked entries in sorted order (duplicates can be appended again).
    for (uint32_t j = lane; j < 128; j += kWave) added[j] = 0;
    wave_sync();
    for (uint32_t a = 0; a < nk; ++a) {
      const uint32_t kid = kept[a];
      uint32_t first = 0xFFFFFFFFu;
      for (uint32_t j0 = 0; j0 < n && first == 0xFFFFFFFFu; j0 += kWave) {
        const uint32_t j = j0 + lane;
        const uint64_t m = __ballot(j < n && sid[uint32_t(skey[j])] == kid);
        if (m) first = j0 + __ffsll((long long)m) - 1;
      }
      if (lane == 0 && first != 0xFFFFFFFFu) added[first] = 1;
      wave_sync();
    }
    if (lane == 0) {
      for (uint32_t i = 0; i < n && nk < deg; ++i) {
        if (added[i]) continue;
        const uint32_t id = sid[uint32_t(skey[i])];
        if (id != kTomb && id != c && id < n_base) kept[nk++] = id;
      }
    }
    nk = __shfl(nk, 0, 64);
    wave_sync();
  }
  for (uint32_t j = lane; j < deg; j += kWave) o[j] = j < nk ? kept[j] : kTomb;
  if (lane == 0) {
    stat_add(stats, ST_MERGE_DIST, n);
    stat_add(stats, ST_MERGE_PAIRS, pairs);
    stat_add(stats, ST_SCREEN_FALLBACK, fbk);
    if (seg > cap) stat_add(stats, ST_REV_OVERFLOW, 1);
  }
}

// ------------------------------------------------------------ launchers

static inline uint32_t grid_for(uint64_t n, uint32_t block) {
  uint64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return uint32_t(g);
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_u32_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, uint64_t n,
                           unsigned long long v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_u64_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_snapshot_stats(const unsigned long long* stats,
                                 unsigned long long* row, hipStream_t s) {
  snapshot_stats_kernel<<<1, 256, 0, s>>>(stats, row);
  return hipGetLastError();
}

hipError_t launch_match(const uint32_t* gt, uint32_t n_query, uint32_t gt_k,
                        uint32_t n_base, uint32_t* winner, hipStream_t s) {
  hipError_t e = launch_fill_u32(winner, n_base, kTomb, s);
  if (e != hipSuccess) return e;
  return launch_match_rows(gt, 0, n_query, gt_k, n_base, winner, s);
}

hipError_t launch_match_rows(const uint32_t* gt, uint32_t q0, uint32_t n_query,
                             uint32_t gt_k, uint32_t n_base, uint32_t* winner,
                             hipStream_t s) {
  if (n_query == 0) return hipSuccess;
  match_winner_kernel<<<grid_for(n_query, 256), 256, 0, s>>>(
      gt, q0, n_query, gt_k, n_base, winner);
  return hipGetLastError();
}

hipError_t launch_prune(const float* base, uint32_t n_base, uint32_t dim,
                        const uint32_t* rows, uint32_t D, const uint32_t* gt,
                        const uint32_t* winner, uint32_t gt_k, uint32_t mp,
                        uint32_t* pruned, uint32_t* pcnt,
                        unsigned long long* stats, const ScreenArgs* sc,
                        bool sweep, const uint32_t* order, hipStream_t s) {
  if (n_base == 0) return hipSuccess;
  const ScreenArgs scv =
      sc ? *sc : ScreenArgs{nullptr, nullptr, nullptr, nullptr, 0, 0.0f, 0.0f, 0.0f};
  if (D > 512) return hipErrorInvalidValue;
  if (rows != nullptr && sweep) {
    // kept-major sweep: screened on the plain BF16 bound; exact FP32 chains
    // for the FP32 path and when handed split-screen arguments (the C ABI
    // hands the sweep plain-screen arguments in every screened context)
#define PG_SWEEP(DP, NC, SCR)                                        \
  prune_sweep_kernel<DP, NC, SCR><<<n_base, 64, 0, s>>>(              \
      base, n_base, dim, rows, D, mp, pruned, pcnt, stats, scv, order)
    if (sc == nullptr || sc->base16lo != nullptr) {
      if (D <= 256) PG_SWEEP(256, 0, false);
      else PG_SWEEP(512, 0, false);
    } else if (sc->dpad == 224) {
      // 7 chunks exactly (dims 193-224): no zero chunk in the registers
      if (D <= 256) PG_SWEEP(256, 7, true);
      else PG_SWEEP(512, 7, true);
    } else if (sc->dpad <= 256) {
      if (D <= 256) PG_SWEEP(256, 8, true);
      else PG_SWEEP(512, 8, true);
    } else if (sc->dpad <= 512) {
      if (D <= 256) PG_SWEEP(256, 16, true);
      else PG_SWEEP(512, 16, true);
    } else {
      return hipErrorInvalidValue;
    }
#undef PG_SWEEP
    return hipGetLastError();
  }
#define PG_PRUNE(DP, NC, NK, SP)                                           \
  prune_kernel<DP, NC, NK, SP><<<n_base, 64, 0, s>>>(                      \
      base, n_base, dim, rows, D, gt, winner, gt_k, mp, pruned, pcnt, stats, \
      scv, rows ? order : nullptr)
#define PG_PRUNE_D(NC, NK, SP)                        \
  do {                                                \
    if (D <= 256) PG_PRUNE(256, NC, NK, SP);          \
    else PG_PRUNE(512, NC, NK, SP);                   \
  } while (0)
#define PG_PRUNE_K(NC, SP)                            \
  do {                                                \
    if (mp <= 32) PG_PRUNE_D(NC, 2, SP);              \
    else if (mp <= 64) PG_PRUNE_D(NC, 4, SP);         \
    else PG_PRUNE_D(NC, 8, SP);                       \
  } while (0)
#define PG_PRUNE_S(NC)                                \
  do {                                                \
    if (scv.base16lo) PG_PRUNE_K(NC, true);           \
    else PG_PRUNE_K(NC, false);                       \
  } while (0)
  if (D > 512) return hipErrorInvalidValue;
  if (sc == nullptr) {
    if (D <= 64) PG_PRUNE(64, 0, 1, false);
    else if (D <= 128) PG_PRUNE(128, 0, 1, false);
    else if (D <= 256) PG_PRUNE(256, 0, 1, false);
    else PG_PRUNE(512, 0, 1, false);
  } else if (sc->dpad <= 256) {
    PG_PRUNE_S(8);
  } else if (sc->dpad <= 512) {
    PG_PRUNE_S(16);
  } else {
    return hipErrorInvalidValue;
  }
#undef PG_PRUNE_S
#undef PG_PRUNE_K
#undef PG_PRUNE_D
#undef PG_PRUNE
  return hipGetLastError();
}

hipError_t launch_reverse(const uint32_t* pruned, const uint32_t* pcnt,
                          uint32_t n_base, uint32_t mp, uint32_t* deg,
                          uint32_t* off, uint32_t* cursor, uint32_t* scan_tmp,
                          uint32_t* srcs, hipStream_t s) {
  hipError_t e = launch_fill_u32(deg, n_base, 0u, s);
  if (e != hipSuccess) return e;
  e = launch_fill_u32(cursor, n_base, 0u, s);
  if (e != hipSuccess) return e;
  const uint64_t total = uint64_t(n_base) * mp;
  if (total == 0) return launch_fill_u32(off, uint64_t(n_base) + 1, 0u, s);
  const uint32_t rb = (n_base + kRevNodes - 1) / kRevNodes;
  reverse_count_kernel<<<rb, 256, 0, s>>>(pruned, pcnt, n_base, mp, deg);
  const uint32_t nb = (n_base + kScanBlock - 1) / kScanBlock;
  scan_blocks_kernel<<<nb, 256, 0, s>>>(deg, n_base, off, scan_tmp);
  scan_sums_kernel<<<1, 1024, 0, s>>>(scan_tmp, nb, off + n_base);
  scan_add_kernel<<<(n_base + 255) / 256, 256, 0, s>>>(off, n_base, scan_tmp);
  reverse_fill_kernel<<<rb, 256, 0, s>>>(pruned, pcnt, n_base, mp, off,
                                         cursor, srcs);
  return hipGetLastError();
}

hipError_t launch_merge(const float* base, uint32_t n_base, uint32_t dim,
                        const uint32_t* off, const uint32_t* srcs,
                        uint32_t cap, const uint32_t* pruned,
                        const uint32_t* pcnt, uint32_t mp, uint32_t deg,
                        int strict, uint32_t* out, unsigned long long* stats,
                        const ScreenArgs* sc, const uint32_t* order,
                        hipStream_t s) {
  if (n_base == 0) return hipSuccess;
  const ScreenArgs scv =
      sc ? *sc : ScreenArgs{nullptr, nullptr, nullptr, nullptr, 0, 0.0f, 0.0f, 0.0f};
#define PG_MERGE(NC, NK, SP)                                                 \
  merge_kernel<NC, NK, SP><<<n_base, 64, 0, s>>>(base, n_base, dim, off,     \
                                                 srcs, cap, pruned, pcnt, mp, \
                                                 deg, strict, out, stats, scv, \
                                                 order)
#define PG_MERGE_K(NC, SP)                    \
  do {                                        \
    if (deg <= 32) PG_MERGE(NC, 2, SP);       \
    else if (deg <= 64) PG_MERGE(NC, 4, SP);  \
    else PG_MERGE(NC, 8, SP);                 \
  } while (0)
#define PG_MERGE_S(NC)                        \
  do {                                        \
    if (scv.base16lo) PG_MERGE_K(NC, true);   \
    else PG_MERGE_K(NC, false);               \
  } while (0)
  if (sc == nullptr || !strict) {
    PG_MERGE(0, 1, false);  // the non-strict fill runs on the plain kernel
  } else if (sc->dpad == 224) {
    PG_MERGE_S(7);  // 7 chunks exactly (dims 193-224)
  } else if (sc->dpad <= 256) {
    PG_MERGE_S(8);
  } else if (sc->dpad <= 512) {
    PG_MERGE_S(16);
  } else {
    return hipErrorInvalidValue;
  }
#undef PG_MERGE_S
#undef PG_MERGE_K
#undef PG_MERGE
  return hipGetLastError();
}


size_t search_smem_bytes(uint32_t table_size, uint32_t lp2, uint32_t kpkd,
                         uint32_t Kp) {
  return ((size_t(table_size) * 4 + 15) & ~size_t(15)) + size_t(lp2) * 8 +
         size_t(kpkd) * 8 + size_t(Kp) * 4 + 16;
}


hipError_t launch_search(const float* queries, uint32_t n_query,
                         const float* base, const void* base16,
                         const void* base8, const float* scale8, uint32_t dpad,
                         uint32_t n_base, uint32_t dim,
                         const uint32_t* graph, uint32_t gdeg, uint32_t Km,
                         uint32_t Kp, uint32_t Kd, uint32_t topk,
                         uint32_t reset_iter, uint32_t entry_seed,
                         uint32_t max_iters, uint32_t table_size,
                         uint32_t gather_batch, uint32_t* out,
                         unsigned long long* stats, uint32_t* queue,
                         const uint32_t* order, uint32_t waves,
                         hipStream_t s) {
  if (n_query == 0) return hipSuccess;
  // gather_batch >= 2 puts two 16-candidate groups in flight (screened)
  const bool two = gather_batch >= 2;
  if (base8 != nullptr)
    return launch_search_fp8(base8, scale8, dim, dpad, two, PG_SEARCH_ARGS);
  if (base16 != nullptr)
    return launch_search_bf16(base16, dim, dpad, two, PG_SEARCH_ARGS);
  const uint32_t tq = (dim + 31) / 32;
  // gather_batch: candidate pairs whose rows are in flight together
  // (perf only); 0 = 4.
  const uint32_t u = gather_batch == 0 ? 4u : gather_batch;
  if (tq <= 4) return launch_search_f32_4(base, dim, u, PG_SEARCH_ARGS);
  if (tq <= 8) return launch_search_f32_8(base, dim, u, PG_SEARCH_ARGS);
  if (tq <= 16) return launch_search_f32_16(base, dim, u, PG_SEARCH_ARGS);
  if (tq <= 32) return launch_search_f32_32(base, dim, u, PG_SEARCH_ARGS);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------- locality order
// Groups nodes by nearest pivot row so that a persistent search grid can
// take queries whose neighbourhoods overlap at the same time (their row
// gathers then hit the XCD's L2 and the MALL instead of HBM).  The order
// never changes a result: every query's search is independent of when it
// runs.

constexpr uint32_t kOrderChunk = 4096;  // nodes per block in count / scatter

// Nearest pivot per node: one thread per node holds its first ds elements;
// the pivots are staged through LDS kTile at a time.  Score |p|^2 - 2 x.p
// (squared L2 minus |x|^2), ties to the lower pivot.
__global__ __launch_bounds__(256) void pivot_group_kernel(
    const float* __restrict__ base, uint32_t n, uint32_t dim, uint32_t ds,
    uint32_t npiv, uint32_t* __restrict__ group) {
  constexpr uint32_t kTile = 128;
  __shared__ float piv[kTile][kOrderDims];
  __shared__ float pn[kTile];
  const uint32_t t = threadIdx.x;
  const uint32_t i = blockIdx.x * 256u + t;
  const float* row = base + size_t(i < n ? i : 0u) * dim;
  float x[kOrderDims];
#pragma unroll
  for (uint32_t d = 0; d < kOrderDims; ++d) x[d] = d < ds ? row[d] : 0.0f;
  float best = FLT_MAX;
  uint32_t bg = 0;
  for (uint32_t p0 = 0; p0 < npiv; p0 += kTile) {
    __syncthreads();
    for (uint32_t e = t; e < kTile * kOrderDims; e += 256u) {
      const uint32_t pp = e / kOrderDims, d = e % kOrderDims;
      const uint32_t g = p0 + pp;
      float v = 0.0f;
      if (g < npiv && d < ds)
        v = base[size_t(uint64_t(g) * n / npiv) * dim + d];
      piv[pp][d] = v;
    }
    __syncthreads();
    if (t < kTile) {
      float a = 0.0f;
      for (uint32_t d = 0; d < kOrderDims; ++d)
        a = __builtin_fmaf(piv[t][d], piv[t][d], a);
      pn[t] = a;
    }
    __syncthreads();
    const uint32_t pe = npiv - p0 < kTile ? npiv - p0 : kTile;
    for (uint32_t pp = 0; pp < pe; ++pp) {
      float dot = 0.0f;
#pragma unroll
      for (uint32_t d = 0; d < kOrderDims; ++d)
        dot = __builtin_fmaf(x[d], piv[pp][d], dot);
      const float sc = pn[pp] - 2.0f * dot;
      if (sc < best) {
        best = sc;
        bg = p0 + pp;
      }
    }
  }
  if (i < n) group[i] = bg;
}

// Group sizes: LDS histogram per block of kOrderChunk nodes, one global
// atomic per non-empty group.
__global__ __launch_bounds__(256) void group_count_kernel(
    const uint32_t* __restrict__ group, uint32_t n, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t h[kOrderGroups];
  const uint32_t t = threadIdx.x;
  h[t] = 0u;
  __syncthreads();
  const uint32_t i0 = blockIdx.x * kOrderChunk;
#pragma unroll
  for (uint32_t k = 0; k < kOrderChunk / 256u; ++k) {
    const uint32_t i = i0 + t + 256u * k;
    if (i < n) atomicAdd(&h[group[i]], 1u);
  }
  __syncthreads();
  if (h[t]) atomicAdd(&cnt[t], h[t]);
}

// Scatter: each block ranks its nodes per group in LDS, reserves one run
// per group with a global atomic (cur), and writes the node ids at
// (group offset + run + rank).  The order inside a group is not fixed.
__global__ __launch_bounds__(256) void group_scatter_kernel(
    const uint32_t* __restrict__ group, uint32_t n,
    const uint32_t* __restrict__ cnt, uint32_t* __restrict__ cur,
    uint32_t* __restrict__ order) {
  __shared__ uint32_t off[kOrderGroups], lc[kOrderGroups], lb[kOrderGroups];
  const uint32_t t = threadIdx.x;
  const uint32_t mine = cnt[t];
  off[t] = mine;
  lc[t] = 0u;
  __syncthreads();
  for (uint32_t d = 1; d < kOrderGroups; d <<= 1) {
    const uint32_t v = t >= d ? off[t - d] : 0u;
    __syncthreads();
    off[t] += v;
    __syncthreads();
  }
  constexpr uint32_t kPer = kOrderChunk / 256u;
  const uint32_t i0 = blockIdx.x * kOrderChunk;
  uint32_t g[kPer], r[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = i0 + t + 256u * k;
    g[k] = 0u;
    r[k] = 0u;
    if (i < n) {
      g[k] = group[i];
      r[k] = atomicAdd(&lc[g[k]], 1u);
    }
  }
  __syncthreads();
  lb[t] = (off[t] - mine) + (lc[t] ? atomicAdd(&cur[t], lc[t]) : 0u);
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = i0 + t + 256u * k;
    if (i < n) order[lb[g[k]] + r[k]] = i;
  }
}

hipError_t launch_locality_order(const float* base, uint32_t n, uint32_t dim,
                                 uint32_t* group, uint32_t* cnt,
                                 uint32_t* order, hipStream_t s) {
  if (n == 0) return hipSuccess;
  static_assert(kOrderGroups == 256, "one thread per group in count/scatter");
  const uint32_t npiv = n < kOrderGroups ? n : kOrderGroups;
  const uint32_t ds = dim < kOrderDims ? dim : kOrderDims;
  hipError_t e = launch_fill_u32(cnt, 2ull * kOrderGroups, 0u, s);
  if (e != hipSuccess) return e;
  pivot_group_kernel<<<(n + 255u) / 256u, 256, 0, s>>>(base, n, dim, ds, npiv,
                                                       group);
  const uint32_t nb = (n + kOrderChunk - 1) / kOrderChunk;
  group_count_kernel<<<nb, 256, 0, s>>>(group, n, cnt);
  group_scatter_kernel<<<nb, 256, 0, s>>>(group, n, cnt, cnt + kOrderGroups,
                                          order);
  return hipGetLastError();
}

// Screened-path preparation, one wave per row: the BF16 copy of the row
// (round to nearest even, zero-padded to dpad), optionally the residual
// copy lo = bf16(x - hi) for the split screen, and the FP32 norm.
__global__ __launch_bounds__(256) void prepare_screen_kernel(
    const float* __restrict__ base, uint32_t n, uint32_t dim, uint32_t dpad,
    __bf16* __restrict__ out16, __bf16* __restrict__ lo16,
    float* __restrict__ norms, float* __restrict__ rnorms) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x / 64u;
  for (uint32_t r = blockIdx.x * wpb + threadIdx.x / 64u; r < n;
       r += gridDim.x * wpb) {
    const float* x = base + size_t(r) * dim;
    float ss = 0.0f, rr = 0.0f;
    for (uint32_t c = lane; c < dpad; c += 64u) {
      const float v = c < dim ? x[c] : 0.0f;
      const __bf16 h = (__bf16)v;
      out16[size_t(r) * dpad + c] = h;
      const float res = v - (float)h;  // exact in FP32
      if (lo16) lo16[size_t(r) * dpad + c] = (__bf16)res;
      ss = __builtin_fmaf(v, v, ss);
      rr = __builtin_fmaf(res, res, rr);
    }
    ss = wave_sum(ss);
    rr = wave_sum(rr);
    if (lane == 0) {
      norms[r] = __builtin_sqrtf(ss);
      if (rnorms) rnorms[r] = __builtin_sqrtf(rr);
    }
  }
}

// FP8 rows for the FP8-screened search, one wave per row: s = max|x|/448
// (1 when the row is zero), bytes e4m3(x / s) zero-padded to dpad.
__global__ __launch_bounds__(256) void prepare_fp8_kernel(
    const float* __restrict__ base, uint32_t n, uint32_t dim, uint32_t dpad,
    uint8_t* __restrict__ out8, float* __restrict__ scale) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x / 64u;
  for (uint32_t r = blockIdx.x * wpb + threadIdx.x / 64u; r < n;
       r += gridDim.x * wpb) {
    const float* x = base + size_t(r) * dim;
    float mx = 0.0f;
    for (uint32_t c = lane; c < dim; c += 64u) mx = fmaxf(mx, fabsf(x[c]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      mx = fmaxf(mx, xor_lane_dyn(mx, uint32_t(o)));
    const float sc = mx > 0.0f ? mx / 448.0f : 1.0f;
    const float inv = 1.0f / sc;
    // 8 consecutive bytes per lane per step
    for (uint32_t c8 = lane * 8u; c8 < dpad; c8 += 512u) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        v[j] = (c8 + j) < dim ? x[c8 + j] * inv : 0.0f;
      *reinterpret_cast<long*>(out8 + size_t(r) * dpad + c8) = pack_fp8x8(v);
    }
    if (lane == 0) scale[r] = sc;
  }
}

hipError_t launch_prepare_fp8(const float* base, uint32_t n, uint32_t dim,
                              uint32_t dpad, void* out8, float* scale,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 3) / 4 < 65536u ? (n + 3) / 4 : 65536u;
  prepare_fp8_kernel<<<blocks, 256, 0, s>>>(
      base, n, dim, dpad, reinterpret_cast<uint8_t*>(out8), scale);
  return hipGetLastError();
}

hipError_t launch_prepare_screen(const float* base, uint32_t n, uint32_t dim,
                                 uint32_t dpad, void* out16, void* lo16,
                                 float* norms, float* rnorms, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 3) / 4 < 65536u ? (n + 3) / 4 : 65536u;
  prepare_screen_kernel<<<blocks, 256, 0, s>>>(
      base, n, dim, dpad, reinterpret_cast<__bf16*>(out16),
      reinterpret_cast<__bf16*>(lo16), norms, rnorms);
  return hipGetLastError();
}

}  // namespace pg

namespace pg {

// Distances of given (query, base id) pairs through the search's own
// evaluators: one one-wave workgroup per query, its k ids staged in LDS,
// distances written as floats (NaN for ids >= n_base).  Used to measure the
// screening error of the BF16 / FP8 paths and the FP32 summation orders
// against FP64.
constexpr uint32_t kPairMaxK = 1024;

template <typename Eval>
__global__ __launch_bounds__(64) void pair_dist_kernel(
    Eval ev, const float* __restrict__ queries, uint32_t n_query,
    const uint32_t* __restrict__ ids, uint32_t k, uint32_t n_base,
    float* __restrict__ out) {
  __shared__ uint32_t todo[kPairMaxK];  // ids of the valid pairs
  __shared__ uint32_t slot[kPairMaxK];  // their positions in the row
  __shared__ uint64_t tail[kPairMaxK];  // their keys, in todo order
  __shared__ uint32_t ntodo_s;
  // query fragments of evaluators that keep them in LDS
  __shared__ bf16x8 qbuf[Eval::kQueryLdsBytes ? Eval::kQueryLdsBytes / 16 : 1];
  ev.qs = qbuf;
  const uint32_t q = blockIdx.x;
  const uint32_t lane = lane_id();
  if (lane == 0) ntodo_s = 0;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < k; t0 += kWave) {
    const uint32_t t = t0 + lane;
    const uint32_t id = t < k ? ids[size_t(q) * k + t] : kTomb;
    const bool ok = t < k && id < n_base;
    if (t < k && !ok) out[size_t(q) * k + t] = __builtin_nanf("");
    const uint64_t bal = __ballot(ok);
    const uint32_t base_pos = ntodo_s;
    __syncthreads();
    if (ok) {
      const uint32_t i = base_pos + __popcll(bal & lanemask_lt());
      todo[i] = id;
      slot[i] = t;
    }
    if (lane == 0) ntodo_s = base_pos + __popcll(bal);
    __syncthreads();
  }
  const uint32_t ntodo = ntodo_s;
  ev.load_query(queries, q);
  __syncthreads();
  ev.eval(todo, ntodo, tail);
  __syncthreads();
  for (uint32_t i = lane; i < ntodo; i += kWave)
    out[size_t(q) * k + slot[i]] = key_dist(tail[i]);
}

// FP32 chain distance of the prune / merge stages for the same pairs.
__global__ __launch_bounds__(256) void pair_chain_kernel(
    const float* __restrict__ queries, uint32_t n_query,
    const float* __restrict__ base, uint32_t dim,
    const uint32_t* __restrict__ ids, uint32_t k, uint32_t n_base,
    float* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= uint64_t(n_query) * k) return;
  const uint32_t q = uint32_t(i / k);
  const uint32_t id = ids[i];
  out[i] = id < n_base ? chain_dist(queries + size_t(q) * dim,
                                    base + size_t(id) * dim, dim)
                       : __builtin_nanf("");
}

template <typename Eval>
static hipError_t launch_pairs_eval(const Eval& ev, const float* queries,
                                    uint32_t n_query, const uint32_t* ids,
                                    uint32_t k, uint32_t n_base, float* out,
                                    hipStream_t s) {
  pair_dist_kernel<Eval><<<n_query, 64, 0, s>>>(ev, queries, n_query, ids, k,
                                                n_base, out);
  return hipGetLastError();
}

template <int TQ>
static hipError_t pairs_f32(const float* base, uint32_t dim,
                            const float* queries, uint32_t n_query,
                            const uint32_t* ids, uint32_t k, uint32_t n_base,
                            float* out, hipStream_t s) {
  EvalF32<TQ, 4> ev;
  ev.base = base;
  ev.dim = dim;
  return launch_pairs_eval(ev, queries, n_query, ids, k, n_base, out, s);
}

template <int NCH>
static hipError_t pairs_bf16(const void* base16, uint32_t dim, uint32_t dpad,
                             const float* queries, uint32_t n_query,
                             const uint32_t* ids, uint32_t k, uint32_t n_base,
                             float* out, hipStream_t s) {
  EvalBF16<NCH, 1> ev;
  ev.base16 = reinterpret_cast<const __bf16*>(base16);
  ev.dim = dim;
  ev.dpad = dpad;
  ev.nch = dpad / 32;
  return launch_pairs_eval(ev, queries, n_query, ids, k, n_base, out, s);
}

template <int NCH>
static hipError_t pairs_fp8(const void* base8, const float* scale8,
                            uint32_t dim, uint32_t dpad, const float* queries,
                            uint32_t n_query, const uint32_t* ids, uint32_t k,
                            uint32_t n_base, float* out, hipStream_t s) {
  EvalFP8<NCH, 1> ev;
  ev.base8 = reinterpret_cast<const uint8_t*>(base8);
  ev.scale8 = scale8;
  ev.dim = dim;
  ev.dpad = dpad;
  ev.nch = dpad / 32;
  return launch_pairs_eval(ev, queries, n_query, ids, k, n_base, out, s);
}

hipError_t launch_pair_dist(const float* queries, uint32_t n_query,
                            const float* base, const void* base16,
                            const void* base8, const float* scale8,
                            uint32_t dpad, uint32_t n_base, uint32_t dim,
                            const uint32_t* ids, uint32_t k, int chain,
                            float* out, hipStream_t s) {
  if (n_query == 0 || k == 0) return hipSuccess;
  if (k > kPairMaxK) return hipErrorInvalidValue;
  if (chain) {
    const uint64_t total = uint64_t(n_query) * k;
    const uint64_t blocks = (total + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    pair_chain_kernel<<<uint32_t(blocks), 256, 0, s>>>(
        queries, n_query, base, dim, ids, k, n_base, out);
    return hipGetLastError();
  }
  const uint32_t nch = dpad / 32;
  if (base8 != nullptr) {
    if (nch <= 8)
      return pairs_fp8<8>(base8, scale8, dim, dpad, queries, n_query, ids, k,
                          n_base, out, s);
    if (nch <= 16)
      return pairs_fp8<16>(base8, scale8, dim, dpad, queries, n_query, ids, k,
                           n_base, out, s);
    return hipErrorInvalidValue;
  }
  if (base16 != nullptr) {
    if (nch <= 8)
      return pairs_bf16<8>(base16, dim, dpad, queries, n_query, ids, k, n_base,
                           out, s);
    if (nch <= 16)
      return pairs_bf16<16>(base16, dim, dpad, queries, n_query, ids, k,
                            n_base, out, s);
    return hipErrorInvalidValue;
  }
  const uint32_t tq = (dim + 31) / 32;
  if (tq <= 4) return pairs_f32<4>(base, dim, queries, n_query, ids, k, n_base, out, s);
  if (tq <= 8) return pairs_f32<8>(base, dim, queries, n_query, ids, k, n_base, out, s);
  if (tq <= 16) return pairs_f32<16>(base, dim, queries, n_query, ids, k, n_base, out, s);
  if (tq <= 32) return pairs_f32<32>(base, dim, queries, n_query, ids, k, n_base, out, s);
  return hipErrorInvalidValue;
}

// ------------------------------------------------- host stages on the GPU
// The reference's CPU routines around the GPU stage -- MatchSup, FusionNN
// and FusionFinal (tests/test_t2iend.cu:365-613) -- as wave-per-list
// kernels under the canonical rules of include/paragraph_host.h (C6-C8), so
// that the lists equal libparagraph_host's id for id.

// The reference CPU distance `compare` (tests/test_t2iend.cu:31-78): 16 FMA
// lanes over blocks of 16, folded to 8 (+ one block of 8) and 4 (+ one
// block of 4), the masked remainder on lanes 1..r (the element past the row
// end taken as 0, rule C8), then ((l0 + l1) + (l2 + l3)).  IEEE FP32 adds
// and FMAs, so one thread's result is bit-equal to the host's.
__device__ __forceinline__ float cpu_order_dist(const float* __restrict__ a,
                                                const float* __restrict__ b,
                                                uint32_t n) {
  float s[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) s[j] = 0.0f;
  uint32_t i = 0;
  for (; n - i >= 16; i += 16) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
      s[j] = __builtin_fmaf(a[i + j], b[i + j], s[j]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = s[8 + j] + s[j];
  if (n - i >= 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = __builtin_fmaf(a[i + j], b[i + j], s[j]);
    i += 8;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) s[j] = s[4 + j] + s[j];
  if (n - i >= 4) {
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = __builtin_fmaf(a[i + j], b[i + j], s[j]);
    i += 4;
  }
  const uint32_t r = n - i;
  if (r > 0) {
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const bool on = j >= 1 && j <= r && i + j < n;
      s[j] = __builtin_fmaf(on ? a[i + j] : 0.0f, on ? b[i + j] : 0.0f, s[j]);
    }
  }
  return -((s[0] + s[1]) + (s[2] + s[3]));
}

constexpr uint32_t kFuseMaxKeep = 128;
constexpr uint32_t kRngStrict = 0;    // RNGPrune, strict
constexpr uint32_t kRngFill = 1;      // RNGPrune, non-strict fill
constexpr uint32_t kRngSmallAsc = 2;  // <= M ids: ascending (C7); else strict

// RNGPrune (tests/test_t2iend.cu:99-177) of one candidate list per wave.
// The list's candidates are cand[l * cstride .. + ccnt[l]); ids >= N, the
// node itself (tgt[l], or l) and repeated ids drop out (the callers' seen
// sets), the rest is sorted by (distance, id).  Kept-major form: the first
// remaining entry is kept and removes every later entry it occludes, which
// keeps exactly the entries the reference's candidate loop keeps (an entry
// survives iff no earlier kept entry occludes it; the cap stops both).  Its
// second pass keeps nothing new.  The non-strict fill appends the other
// entries in sorted order.  W: candidate slots (128 or 256).
template <uint32_t W>
__global__ __launch_bounds__(64) void fuse_rng_kernel(
    const float* __restrict__ base, uint32_t n_base, uint32_t dim,
    const uint32_t* __restrict__ cand, uint32_t cstride,
    const uint32_t* __restrict__ ccnt, const uint32_t* __restrict__ tgt,
    uint32_t M, uint32_t mode, uint32_t* __restrict__ out, uint32_t ostride,
    uint32_t* __restrict__ ocnt) {
  __shared__ uint32_t ids[W];
  __shared__ uint64_t keys[W];
  __shared__ float dv[W];
  __shared__ uint8_t alive[W];
  __shared__ uint8_t keptf[W];
  __shared__ uint32_t kept[kFuseMaxKeep];
  constexpr uint32_t S = W / kWave;
  const uint32_t l = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t t = tgt ? tgt[l] : l;
  uint32_t n = ccnt[l];
  if (n > W) n = W;
  if (t >= n_base) n = 0;  // a list without a node (an unmatched query)
  for (uint32_t j = lane; j < W; j += kWave) {
    uint32_t id = kTomb;
    if (j < n) {
      id = cand[size_t(l) * cstride + j];
      if (id >= n_base || id == t) id = kTomb;
    }
    ids[j] = id;
  }
  __syncthreads();
  sort_list(ids, W);
  // distinct ids, ascending, compacted
  uint32_t v[S];
  bool keep[S];
#pragma unroll
  for (uint32_t sl = 0; sl < S; ++sl) {
    const uint32_t j = lane + kWave * sl;
    v[sl] = ids[j];
    keep[sl] = v[sl] != kTomb && (j == 0 || ids[j - 1] != v[sl]);
  }
  __syncthreads();
  uint32_t m = 0;
#pragma unroll
  for (uint32_t sl = 0; sl < S; ++sl) {
    const uint64_t km = __ballot(keep[sl]);
    if (keep[sl]) ids[m + __popcll(km & lanemask_lt())] = v[sl];
    m += __popcll(km);
  }
  __syncthreads();
  uint32_t* o = out + size_t(l) * ostride;
  if (mode == kRngSmallAsc && m <= M) {
    for (uint32_t j = lane; j < m; j += kWave) o[j] = ids[j];
    if (lane == 0) ocnt[l] = m;
    return;
  }
  const float* xt = base + size_t(t < n_base ? t : 0u) * dim;
  for (uint32_t j = lane; j < W; j += kWave) {
    uint64_t k = kPadKey;
    if (j < m) {
      const uint32_t id = ids[j];
      const float d = cpu_order_dist(xt, base + size_t(id) * dim, dim);
      k = (uint64_t(ordered_bits(d)) << 32) | id;
    }
    keys[j] = k;
  }
  __syncthreads();
  sort_list(keys, W);
  for (uint32_t j = lane; j < W; j += kWave) {
    alive[j] = j < m ? 1 : 0;
    keptf[j] = 0;
    dv[j] = j < m ? from_ordered(uint32_t(keys[j] >> 32)) : FLT_MAX;
  }
  __syncthreads();
  uint32_t nk = 0, pos = 0;
  while (nk < M) {
    uint32_t f = W;
    for (uint32_t j0 = pos & ~(kWave - 1u); j0 < m && f == W; j0 += kWave) {
      const uint32_t j = j0 + lane;
      const uint64_t am = __ballot(j >= pos && j < m && alive[j]);
      if (am) f = j0 + __ffsll((long long)am) - 1;
    }
    if (f == W) break;
    const uint32_t kid = uint32_t(keys[f]);
    if (lane == 0) {
      kept[nk] = kid;
      keptf[f] = 1;
    }
    ++nk;
    pos = f + 1;
    if (nk == M) break;
    const float* xk = base + size_t(kid) * dim;
    for (uint32_t j = pos + lane; j < m; j += kWave) {
      if (alive[j] &&
          cpu_order_dist(base + size_t(uint32_t(keys[j])) * dim, xk, dim) <
              dv[j])
        alive[j] = 0;
    }
    __syncthreads();
  }
  __syncthreads();
  if (mode == kRngFill && nk < M) {
    for (uint32_t j0 = 0; j0 < m && nk < M; j0 += kWave) {
      const uint32_t j = j0 + lane;
      const bool add = j < m && !keptf[j];
      const uint64_t am = __ballot(add);
      const uint32_t r = __popcll(am & lanemask_lt());
      if (add && nk + r < M) kept[nk + r] = uint32_t(keys[j]);
      const uint32_t c = __popcll(am);
      nk = nk + c < M ? nk + c : M;
    }
    __syncthreads();
  }
  for (uint32_t j = lane; j < nk; j += kWave) o[j] = kept[j];
  if (lane == 0) ocnt[l] = nk;
}

// MatchSup's first-come match on gt column `col` (tests/test_t2iend.cu:
// 376-390): the lowest query id wins a base.  owner must be all 1s before.
__global__ void column_owner_kernel(const uint32_t* __restrict__ gt,
                                    uint32_t Q, uint32_t K, uint32_t col,
                                    uint32_t n_base,
                                    uint32_t* __restrict__ owner) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const uint32_t b = gt[size_t(q) * K + col];
  if (b < n_base) atomicMin(&owner[b], q);
}

// MatchSup forward lists (:405-428): list q belongs to the base it matched,
// and its candidates are the gt row up to the first id >= N.
__global__ void sup_forward_lists_kernel(const uint32_t* __restrict__ gt,
                                         uint32_t Q, uint32_t K, uint32_t col,
                                         uint32_t n_base,
                                         const uint32_t* __restrict__ owner,
                                         uint32_t* __restrict__ ccnt,
                                         uint32_t* __restrict__ tgt) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const uint32_t* g = gt + size_t(q) * K;
  const uint32_t b = g[col];
  uint32_t c = 0;
  if (b < n_base && owner[b] == q)
    while (c < K && g[c] < n_base) ++c;
  ccnt[q] = c;
  tgt[q] = c ? b : kTomb;
}

// MatchSup merge candidates (:450-461), one wave per base b: the forward
// list of b, then the sources of the cap lowest query ids among the forward
// lists that keep b (rule C6; the reverse CSR's sources are query ids).
__global__ __launch_bounds__(64) void sup_merge_lists_kernel(
    const uint32_t* __restrict__ gt, uint32_t K, uint32_t col,
    const uint32_t* __restrict__ owner, const uint32_t* __restrict__ fwd,
    const uint32_t* __restrict__ fcnt, uint32_t M,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ srcs,
    uint32_t cap, uint32_t n_base, uint32_t* __restrict__ cand,
    uint32_t cstride, uint32_t* __restrict__ ccnt) {
  __shared__ uint32_t buf[2 * kSegChunk];
  const uint32_t b = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t s0 = off[b], seg = off[b + 1] - s0;
  const uint32_t r = seg < cap ? seg : cap;
  if (seg <= kSegChunk) {
    uint32_t n2 = 64;
    while (n2 < seg) n2 <<= 1;
    for (uint32_t j = lane; j < n2; j += kWave)
      buf[j] = j < seg ? srcs[s0 + j] : kTomb;
    __syncthreads();
    if (seg > 1) sort_list(buf, n2);
  } else {
    for (uint32_t j = lane; j < kSegChunk; j += kWave) buf[j] = srcs[s0 + j];
    __syncthreads();
    bitonic_sort(buf, kSegChunk);
    for (uint32_t done = kSegChunk; done < seg; done += kSegChunk) {
      for (uint32_t j = lane; j < kSegChunk; j += kWave) {
        if (j >= cap) buf[j] = kTomb;
        const uint32_t g = done + j;
        buf[kSegChunk + j] = g < seg ? srcs[s0 + g] : kTomb;
      }
      __syncthreads();
      bitonic_sort(buf, 2 * kSegChunk);
    }
  }
  const uint32_t q0 = owner[b];
  const uint32_t nf = q0 < 0xFFFFFFFFu ? fcnt[q0] : 0u;
  uint32_t* c = cand + size_t(b) * cstride;
  for (uint32_t j = lane; j < nf; j += kWave) c[j] = fwd[size_t(q0) * M + j];
  for (uint32_t j = lane; j < r; j += kWave)
    c[nf + j] = gt[size_t(buf[j]) * K + col];
  if (lane == 0) ccnt[b] = nf + r;
}

// Concatenates per-node rows of up to `w` ids from `parts` row sets (row
// set p at src + p * pstride, counts at cnt + p * n) into cand rows.
__global__ void concat_rows_kernel(const uint32_t* __restrict__ src,
                                   const uint32_t* __restrict__ cnt,
                                   uint32_t parts, uint64_t pstride,
                                   uint32_t w, uint32_t n,
                                   uint32_t* __restrict__ cand,
                                   uint32_t cstride,
                                   uint32_t* __restrict__ ccnt) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  uint32_t k = 0;
  uint32_t* c = cand + size_t(b) * cstride;
  for (uint32_t p = 0; p < parts; ++p) {
    const uint32_t nc = cnt[size_t(p) * n + b];
    const uint32_t* row = src + p * pstride + size_t(b) * w;
    for (uint32_t j = 0; j < nc; ++j) c[k++] = row[j];
  }
  ccnt[b] = k;
}

// FusionFinal candidates (:553-582): the bipartite list, the link row and
// the projection row of each base (ids >= N, the node and repeats drop out
// in fuse_rng_kernel; the node itself, which the reference admits from the
// link row, is never kept and occludes nothing, so dropping it changes no
// list).
__global__ void final_lists_kernel(const uint32_t* __restrict__ bip,
                                   const uint32_t* __restrict__ bcnt,
                                   uint32_t wb,
                                   const uint32_t* __restrict__ link,
                                   uint32_t M_link,
                                   const uint32_t* __restrict__ proj,
                                   uint32_t proj_deg, uint32_t n,
                                   uint32_t* __restrict__ cand,
                                   uint32_t cstride,
                                   uint32_t* __restrict__ ccnt) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  uint32_t* c = cand + size_t(b) * cstride;
  uint32_t k = 0;
  for (uint32_t j = 0; j < bcnt[b]; ++j) c[k++] = bip[size_t(b) * wb + j];
  for (uint32_t j = 0; j < M_link; ++j) c[k++] = link[size_t(b) * M_link + j];
  for (uint32_t j = 0; j < proj_deg; ++j) c[k++] = proj[size_t(b) * proj_deg + j];
  ccnt[b] = k;
}

// FusionFinal's output row (:553-603): the bipartite entries (first
// occurrences, not the node, < N) in their order, then the pruned entries
// not among them, while fewer than M_final - |bipartite| were added; that
// room is unsigned in the reference, so a bipartite list longer than
// M_final takes every pruned entry.
__global__ void final_compose_kernel(const uint32_t* __restrict__ bip,
                                     const uint32_t* __restrict__ bcnt,
                                     uint32_t wb,
                                     const uint32_t* __restrict__ pr,
                                     const uint32_t* __restrict__ prc,
                                     uint32_t M_final, uint32_t n_base,
                                     uint32_t* __restrict__ rows,
                                     uint32_t wr, uint32_t* __restrict__ rcnt) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_base) return;
  uint32_t* row = rows + size_t(b) * wr;
  uint32_t nr = 0;
  for (uint32_t j = 0; j < bcnt[b]; ++j) {
    const uint32_t id = bip[size_t(b) * wb + j];
    if (id >= n_base || id == b) continue;
    bool dup = false;
    for (uint32_t i = 0; i < nr && !dup; ++i) dup = row[i] == id;
    if (!dup) row[nr++] = id;
  }
  const uint32_t nb = nr;
  const uint64_t room =
      nb <= M_final ? uint64_t(M_final - nb) : ~uint64_t(0);
  uint64_t added = 0;
  for (uint32_t j = 0; j < prc[b] && added < room; ++j) {
    const uint32_t id = pr[size_t(b) * M_final + j];
    bool in = false;
    for (uint32_t i = 0; i < nb && !in; ++i) in = row[i] == id;
    if (!in) {
      row[nr++] = id;
      ++added;
    }
  }
  rcnt[b] = nr;
}

// Fixed-stride rows -> CSR ids (offsets from an exclusive scan of the counts).
__global__ void rows_to_csr_kernel(const uint32_t* __restrict__ rows,
                                   uint32_t wr,
                                   const uint32_t* __restrict__ off,
                                   uint32_t n, uint32_t* __restrict__ ids) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const uint32_t o = off[b], c = off[b + 1] - o;
  for (uint32_t j = 0; j < c; ++j) ids[o + j] = rows[size_t(b) * wr + j];
}

static hipError_t launch_fuse_rng(const float* base, uint32_t n_base,
                                  uint32_t dim, const uint32_t* cand,
                                  uint32_t cstride, const uint32_t* ccnt,
                                  const uint32_t* tgt, uint32_t nl,
                                  uint32_t W, uint32_t M, uint32_t mode,
                                  uint32_t* out, uint32_t ostride,
                                  uint32_t* ocnt, hipStream_t s) {
  if (nl == 0) return hipSuccess;
  if (M == 0 || M > kFuseMaxKeep || M > ostride) return hipErrorInvalidValue;
  if (W <= 128)
    fuse_rng_kernel<128><<<nl, 64, 0, s>>>(base, n_base, dim, cand, cstride,
                                           ccnt, tgt, M, mode, out, ostride,
                                           ocnt);
  else if (W <= 256)
    fuse_rng_kernel<256><<<nl, 64, 0, s>>>(base, n_base, dim, cand, cstride,
                                           ccnt, tgt, M, mode, out, ostride,
                                           ocnt);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Workspace of launch_fusion, carved in this order from one buffer.
struct FusionWs {
  uint32_t *owner, *ccnt, *tgt, *fcnt, *deg, *cursor, *off, *bcnt, *scan_tmp;
  uint32_t *fwd, *srcs, *cand, *sup, *scnt, *bip, *pr, *prc, *rows, *rcnt;
  uint64_t words;
};

static FusionWs fusion_carve(const FusionShape& f, uint32_t* base_ptr) {
  const uint64_t N = f.n_base, M = f.M_nn, R = f.rounds;
  const uint64_t nb = (N + kScanBlock - 1) / kScanBlock + 1;
  FusionWs w{};
  uint64_t at = 0;
  auto take = [&](uint64_t words) {
    uint32_t* q = base_ptr ? base_ptr + at : nullptr;
    at += words;
    return q;
  };
  w.owner = take(N);
  w.ccnt = take(N);
  w.tgt = take(N);
  w.fcnt = take(N);
  w.deg = take(N);
  w.cursor = take(N);
  w.off = take(N + 1);
  w.bcnt = take(N);
  w.scan_tmp = take(nb);
  w.fwd = take(N * M);
  w.srcs = take(N * M);
  w.cand = take(N * 256);
  w.sup = take(R * N * M);
  w.scnt = take(R * N);
  w.bip = take(N * 2 * M);
  w.pr = take(N * f.M_final);
  w.prc = take(N);
  w.rows = take(N * (2 * M + f.M_final));
  w.rcnt = take(N);
  w.words = at;
  return w;
}

uint64_t fusion_ws_words(const FusionShape& f) {
  return fusion_carve(f, nullptr).words;
}

#define PG_FUSION_WS                                                        \
  const uint32_t N = f.n_base, K = f.gt_k, M = f.M_nn, R = f.rounds;        \
  const uint32_t dim = f.dim, cap = 128u - M, wf = 2 * M + f.M_final;        \
  const uint32_t nblk = (N + kScanBlock - 1) / kScanBlock;                  \
  const FusionWs w = fusion_carve(f, ws);                                    \
  uint32_t *owner = w.owner, *ccnt = w.ccnt, *tgt = w.tgt, *fcnt = w.fcnt;  \
  uint32_t *deg = w.deg, *cursor = w.cursor, *off = w.off, *bcnt = w.bcnt;  \
  uint32_t *scan_tmp = w.scan_tmp, *fwd = w.fwd, *srcs = w.srcs;            \
  uint32_t *cand = w.cand, *sup = w.sup, *scnt = w.scnt, *bip = w.bip;      \
  uint32_t *pr = w.pr, *prc = w.prc, *rows = w.rows, *rcnt = w.rcnt;        \
  const uint32_t tb = 256, tg = (N + tb - 1) / tb;                          \
  (void)K; (void)cap; (void)wf; (void)nblk; (void)owner; (void)tgt;          \
  (void)fcnt; (void)deg; (void)cursor; (void)off; (void)fwd; (void)srcs;     \
  (void)sup; (void)scnt; (void)pr; (void)prc; (void)rows; (void)rcnt;        \
  (void)R; (void)dim; (void)scan_tmp; (void)cand; (void)ccnt

// MatchSup rounds 1..R and FusionNN: the bipartite rows (bip, bcnt) in the
// workspace.  Reads only base and gt, so it can run beside the GPU stage.
hipError_t launch_fusion_bipartite(const FusionShape& f, const float* base,
                                   const uint32_t* gt, uint32_t* ws,
                                   hipStream_t s) {
  if (f.n_base == 0) return hipSuccess;
  PG_FUSION_WS;
  hipError_t e = hipSuccess;
  for (uint32_t r = 0; r < R && e == hipSuccess; ++r) {
    const uint32_t col = r + 1;
    e = launch_fill_u32(owner, N, 0xFFFFFFFFu, s);
    if (e != hipSuccess) break;
    column_owner_kernel<<<tg, tb, 0, s>>>(gt, N, K, col, N, owner);
    sup_forward_lists_kernel<<<tg, tb, 0, s>>>(gt, N, K, col, N, owner, ccnt,
                                               tgt);
    e = launch_fuse_rng(base, N, dim, gt, K, ccnt, tgt, N, K, M, kRngFill,
                        fwd, M, fcnt, s);
    if (e != hipSuccess) break;
    // reverse lists keyed by base, sources = query ids (rows of fwd)
    e = launch_reverse(fwd, fcnt, N, M, deg, off, cursor, scan_tmp, srcs, s);
    if (e != hipSuccess) break;
    sup_merge_lists_kernel<<<N, 64, 0, s>>>(gt, K, col, owner, fwd, fcnt, M,
                                            off, srcs, cap, N, cand, 128,
                                            ccnt);
    e = launch_fuse_rng(base, N, dim, cand, 128, ccnt, nullptr, N, 128, M,
                        kRngSmallAsc, sup + uint64_t(r) * N * M, M,
                        scnt + uint64_t(r) * N, s);
  }
  if (e != hipSuccess) return e;
  // FusionNN: union of the supply lists, strict to 2 M
  concat_rows_kernel<<<tg, tb, 0, s>>>(sup, scnt, R, uint64_t(N) * M, M, N,
                                       cand, 256, ccnt);
  e = launch_fuse_rng(base, N, dim, cand, 256, ccnt, nullptr, N, R * M,
                      2 * M, kRngStrict, bip, 2 * M, bcnt, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// FusionFinal from the workspace's bipartite rows, the link rows and the
// projection rows, as CSR.
hipError_t launch_fusion_final(const FusionShape& f, const float* base,
                               const uint32_t* proj, const uint32_t* link,
                               uint32_t* ws, uint32_t* off_out,
                               uint32_t* ids_out, hipStream_t s) {
  if (f.n_base == 0) return hipSuccess;
  PG_FUSION_WS;
  final_lists_kernel<<<tg, tb, 0, s>>>(bip, bcnt, 2 * M, link, f.M_link, proj,
                                       f.proj_deg, N, cand, 256, ccnt);
  hipError_t e = launch_fuse_rng(base, N, dim, cand, 256, ccnt, nullptr, N,
                                 2 * M + f.M_link + f.proj_deg, f.M_final,
                                 kRngStrict, pr, f.M_final, prc, s);
  if (e != hipSuccess) return e;
  final_compose_kernel<<<tg, tb, 0, s>>>(bip, bcnt, 2 * M, pr, prc, f.M_final,
                                         N, rows, wf, rcnt);
  scan_blocks_kernel<<<nblk, 256, 0, s>>>(rcnt, N, off_out, scan_tmp);
  scan_sums_kernel<<<1, 1024, 0, s>>>(scan_tmp, nblk, off_out + N);
  scan_add_kernel<<<tg, tb, 0, s>>>(off_out, N, scan_tmp);
  rows_to_csr_kernel<<<tg, tb, 0, s>>>(rows, wf, off_out, N, ids_out);
  return hipGetLastError();
}
#undef PG_FUSION_WS

hipError_t launch_fusion(const FusionShape& f, const float* base,
                         const uint32_t* gt, const uint32_t* proj,
                         const uint32_t* link, uint32_t* ws, uint32_t* off_out,
                         uint32_t* ids_out, hipStream_t s) {
  hipError_t e = launch_fusion_bipartite(f, base, gt, ws, s);
  if (e != hipSuccess) return e;
  return launch_fusion_final(f, base, proj, link, ws, off_out, ids_out, s);
}

}  // namespace pg
