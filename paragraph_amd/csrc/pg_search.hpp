// K4 (the search) for MI355X: visited set, distance evaluators, list
// merges, search_kernel and its launch templates.  Each evaluator family is
// instantiated in a translation unit of its own (pg_search_*.hip), so that
// their kernels compile in parallel; pg_kernels.hip dispatches to them.
#pragma once
#include "pg_device.hpp"
#include "pg_kernels.h"

namespace pg {

// ------------------------------------------------------------- K4: search
// link_process_v0_store_base_data (Gsearcher.cuh:896-1066), one wave per
// query.  LDS per wave: the visited set (open addressing, exact), the list
// as u64 keys (distance, id, expanded bit), the compacted miss ids (todo)
// beside a scratch region of the same size (together they hold the merge's
// sorted or entering miss keys), and the Kp expanded ids.  Distances: half a wave (32
// lanes) per candidate, lane l accumulating elements l, l+32, ... with the
// query slice held in registers, then the xor tree 16/8/4/2/1, whose lane-0
// result equals the reference's __shfl_down_sync tree.
// Visited set: open addressing with linear probing over `tsize` slots (any
// size; the host sizes it above the most ids one reset period can insert,
// Km + reset_iter*Kp*Kd, so a probe always ends).  Slot 0 = empty, else
// id + 1.  Multiplicative hash, range-reduced by multiply-shift.
__device__ __forceinline__ uint32_t visit_hash(uint32_t id, uint32_t tsize) {
  return uint32_t((uint64_t(id * 2654435761u) * tsize) >> 32);
}

__device__ __forceinline__ bool visit_insert(uint32_t* tab, uint32_t tsize,
                                             uint32_t id) {
  const uint32_t key = id + 1u;
  uint32_t s = visit_hash(id, tsize);
  while (true) {
    const uint32_t old = atomicCAS(&tab[s], 0u, key);
    if (old == 0u) return true;
    if (old == key) return false;
    s = (s + 1u == tsize) ? 0u : s + 1u;
  }
}

// Distance evaluators for the search.  Each one computes the distances of
// the todo-listed candidates to the query and writes their list keys.
//
// EvalF32: the reference's exact order.  Half a wave (32 lanes) per
// candidate; lane l holds query elements l, l+32, ... in registers and runs
// one FMA chain over them, then the xor tree 16/8/4/2/1 (lane 0 of each half
// ends with exactly the reference's __shfl_down_sync result).  U candidate
// pairs have their rows in flight together.
template <int TQ, int U>
struct EvalF32 {
  static constexpr bool kExactVisited = true;
  static constexpr int kMinWaves = 1;  // occupancy target for the compiler
  static constexpr int kLeanMinWaves = 1;  // the same for the lean variant
  static constexpr uint32_t kQueryLdsBytes = 0;
  bf16x8* qs;  // unused
  float qv[TQ];
  const float* __restrict__ base;
  uint32_t dim;

  __device__ __forceinline__ void load_query(const float* __restrict__ queries,
                                             uint32_t q) {
    load_query_regs(queries, q);
  }
  // the part of the query every wave that evaluates needs (registers)
  __device__ __forceinline__ void load_query_regs(
      const float* __restrict__ queries, uint32_t q) {
    const uint32_t l = lane_id() & 31u;
#pragma unroll
    for (int t = 0; t < TQ; ++t) {
      const uint32_t e = l + 32u * t;
      qv[t] = e < dim ? queries[size_t(q) * dim + e] : 0.0f;
    }
  }

  // Candidate ids todo[0, ntodo) split over `parts` waves: wave `part`
  // takes the candidate pairs p0 = 2U * (part + parts * i); the key of
  // todo[i] goes to tail[i].
  __device__ __forceinline__ void eval(const uint32_t* todo, uint32_t ntodo,
                                       uint64_t* tail, uint32_t part = 0,
                                       uint32_t parts = 1) const {
    const uint32_t lane = lane_id();
    const uint32_t half = lane >> 5, l = lane & 31u;
    for (uint32_t p0 = 2 * U * part; p0 < ntodo; p0 += 2 * U * parts) {
      uint32_t cid[U];
      bool ok[U];
      float xv[U][TQ];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t idx = p0 + 2 * u + half;
        ok[u] = idx < ntodo;
        cid[u] = ok[u] ? todo[idx] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float* row = base + size_t(cid[u]) * dim;
#pragma unroll
        for (int t = 0; t < TQ; ++t) {
          const uint32_t e = l + 32u * t;
          xv[u][t] = (ok[u] && e < dim) ? row[e] : 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < TQ; ++t) {
          if (l + 32u * t < dim) s = __builtin_fmaf(qv[t], xv[u][t], s);
        }
        s = s + xor_lane<16>(s);
        s = s + xor_lane<8>(s);
        s = s + xor_lane<4>(s);
        s = s + xor_lane<2>(s);
        s = s + xor_lane<1>(s);
        if (l == 0 && ok[u]) tail[p0 + 2 * u + half] = search_key(-s, cid[u]);
      }
    }
  }
};

// EvalBF16: screening distances on MFMA.  The base is a BF16 copy with rows
// zero-padded to dpad (a multiple of 32).  16 candidates form the A operand
// of v_mfma_f32_16x16x32_bf16 (lane l holds row l%16, elements
// 32c + 8(l/16) .. +7 of chunk c); the query, broadcast to all 16 columns,
// is the B operand; FP32 accumulation over NCH chunks.  Lane l of the result
// holds rows 4(l/16) .. +3.  G groups of 16 are in flight together.
template <int NCH, int G>
struct EvalBF16 {
  static constexpr bool kExactVisited = false;
  // 6 waves per SIMD (<= 80 VGPRs, a few spilled) for one group of <= 256
  // padded elements: measured 3% faster than 5 waves on the 1M x 200
  // search; 7 waves (72 VGPRs) spill more and gain nothing.  Up to 512
  // padded elements: 4 waves (<= 128 VGPRs, no spills; 3 untargeted)
  static constexpr int kMinWaves =
      (NCH <= 8 && G == 1) ? 6 : (NCH <= 16 && G == 1) ? 4 : 1;
  // the lean variant holds fewer live values: 7 waves (<= 72 VGPRs)
  static constexpr int kLeanMinWaves = (NCH <= 8 && G == 1) ? 7 : kMinWaves;
  // The query's B-operand fragments live in LDS (qs[4c + k] = elements
  // 32c + 8k .. +7), read per MFMA: replicating them in registers would
  // cost 4*NCH VGPRs per lane.
  static constexpr uint32_t kQueryLdsBytes = NCH * 4 * 16;
  bf16x8* qs;
  const __bf16* __restrict__ base16;
  uint32_t dim, dpad, nch;

  __device__ __forceinline__ void load_query(const float* __restrict__ queries,
                                             uint32_t q) {
    const uint32_t lane = lane_id();
    const float* x = queries + size_t(q) * dim;
    if (lane < 4u * NCH) {
      const uint32_t c = lane >> 2, k = lane & 3u;
      float v[8];
      load_query_octet(x, dim, c < nch, 32u * c + 8u * k, v);
      bf16x8 f;
#pragma unroll
      for (int j = 0; j < 8; ++j) f[j] = (__bf16)v[j];
      qs[lane] = f;
    }
  }
  // the query lives in LDS only
  __device__ __forceinline__ void load_query_regs(const float* __restrict__,
                                                  uint32_t) {}

  // Groups of 16 candidate ids split over `parts` waves: wave `part` takes
  // the groups g0 = 16G * (part + parts * i); the key of todo[i] goes to
  // tail[i].
  __device__ __forceinline__ void eval(const uint32_t* todo, uint32_t ntodo,
                                       uint64_t* tail, uint32_t part = 0,
                                       uint32_t parts = 1) const {
    const uint32_t lane = lane_id();
    const uint32_t m = lane & 15u, kq = lane >> 4;
    for (uint32_t g0 = 16u * G * part; g0 < ntodo; g0 += 16u * G * parts) {
      bf16x8 a[G][NCH];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint32_t idx = g0 + 16u * g + m;
        const bool ok = idx < ntodo;
        const uint32_t cid = ok ? todo[idx] : 0u;
        const bf16x8* row = reinterpret_cast<const bf16x8*>(
            base16 + size_t(cid) * dpad + 8u * kq);
        // lanes past ntodo read row 0 (a valid row): their output rows are
        // never stored, so only the chunk bound (uniform) decides a load
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          if (uint32_t(c) < nch) {
            a[g][c] = row[4 * c];
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) a[g][c][j] = (__bf16)0.0f;
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        // keep the fragment reads inside the loop (not hoisted to registers)
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g][c], qs[4 * c + kq],
                                                        acc, 0, 0, 0);
        // lane l (l%16 < 4) writes candidate 4(l/16) + l%4
        if (m < 4u) {
          const uint32_t cm = 4u * kq + m;
          const uint32_t idx = g0 + 16u * g + cm;
          // element m of the accumulator, by selects (no branches)
          float dot = acc[0];
          dot = m == 1u ? acc[1] : dot;
          dot = m == 2u ? acc[2] : dot;
          dot = m == 3u ? acc[3] : dot;
          if (idx < ntodo) tail[idx] = search_key(-dot, todo[idx]);
        }
      }
    }
  }
};

// EvalFP8: screening distances on MFMA in FP8 e4m3 (OCP e4m3fn on gfx950,
// max 448) with one scale per vector: row x is stored as e4m3(x / s_x),
// s_x = max|x| / 448, and the query likewise; v_mfma_f32_16x16x32_fp8_fp8
// accumulates in FP32 and the distance is -(s_q s_x) * acc.  Same operand
// layout as EvalBF16 (lane l: row l%16, bytes 32c + 8(l/16) .. +7).
__device__ __forceinline__ long pack_fp8x8(const float* v) {
  int lo = 0, hi = 0;
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], lo, false);
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], hi, false);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
  return long(uint32_t(lo)) | (long(uint32_t(hi)) << 32);
}

template <int NCH, int G>
struct EvalFP8 {
  static constexpr bool kExactVisited = false;
  static constexpr int kMinWaves = (NCH <= 8 && G == 1) ? 5 : 1;
  static constexpr int kLeanMinWaves = kMinWaves;
  // e4m3 query fragments in LDS: 8 bytes per (chunk, lane group)
  static constexpr uint32_t kQueryLdsBytes = NCH * 4 * 8;
  bf16x8* qs;
  float qscale;
  const uint8_t* __restrict__ base8;
  const float* __restrict__ scale8;
  uint32_t dim, dpad, nch;

  // the query's scale (registers): the largest |element| / 448
  __device__ __forceinline__ void load_query_regs(
      const float* __restrict__ queries, uint32_t q) {
    const float* x = queries + size_t(q) * dim;
    float mx = 0.0f;
    for (uint32_t e = lane_id(); e < dim; e += kWave)
      mx = fmaxf(mx, fabsf(x[e]));
    for (int o = 32; o > 0; o >>= 1)
      mx = fmaxf(mx, xor_lane_dyn(mx, uint32_t(o)));
    qscale = mx > 0.0f ? mx / 448.0f : 1.0f;
  }

  __device__ __forceinline__ void load_query(const float* __restrict__ queries,
                                             uint32_t q) {
    const uint32_t lane = lane_id();
    const float* x = queries + size_t(q) * dim;
    float mx = 0.0f;
    for (uint32_t e = lane; e < dim; e += kWave) mx = fmaxf(mx, fabsf(x[e]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      mx = fmaxf(mx, xor_lane_dyn(mx, uint32_t(o)));
    qscale = mx > 0.0f ? mx / 448.0f : 1.0f;
    const float inv = 1.0f / qscale;
    if (lane < 4u * NCH) {
      const uint32_t c = lane >> 2, k = lane & 3u;
      float v[8];
      load_query_octet(x, dim, c < nch, 32u * c + 8u * k, v);
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] *= inv;
      reinterpret_cast<long*>(qs)[lane] = pack_fp8x8(v);
    }
  }

  // Same split and outputs as EvalBF16::eval.
  __device__ __forceinline__ void eval(const uint32_t* todo, uint32_t ntodo,
                                       uint64_t* tail, uint32_t part = 0,
                                       uint32_t parts = 1) const {
    const uint32_t lane = lane_id();
    const uint32_t m = lane & 15u, kq = lane >> 4;
    for (uint32_t g0 = 16u * G * part; g0 < ntodo; g0 += 16u * G * parts) {
      long a[G][NCH];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const uint32_t idx = g0 + 16u * g + m;
        const bool ok = idx < ntodo;
        const uint32_t cid = ok ? todo[idx] : 0u;
        const long* row = reinterpret_cast<const long*>(
            base8 + size_t(cid) * dpad + 8u * kq);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          // lanes past ntodo read row 0, whose output rows are never stored
          a[g][c] = uint32_t(c) < nch ? row[4 * c] : 0l;
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const long* q8 = reinterpret_cast<const long*>(qs);
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(
              a[g][c], q8[4 * c + kq], acc, 0, 0, 0);
        if (m < 4u) {
          const uint32_t cm = 4u * kq + m;
          const uint32_t idx = g0 + 16u * g + cm;
          // element m of the accumulator, by selects (no branches)
          float dot = acc[0];
          dot = m == 1u ? acc[1] : dot;
          dot = m == 2u ? acc[2] : dot;
          dot = m == 3u ? acc[3] : dot;
          if (idx < ntodo) {
            const uint32_t cid = todo[idx];
            tail[idx] = search_key(-(dot * qscale * scale8[cid]), cid);
          }
        }
      }
    }
  }
};

// Merges the misses of one iteration (keys tail[0, m)) into the top-Km list
// (the reference re-sorts the whole list, Gsearcher.cuh:1045; only its
// first Km entries survive into the next iteration, so a merge is
// equivalent).
// Returns true when this iteration is the last one (no unexpanded entry
// would stay in the top Km, or the watchdog stops after it); the caller then
// sorts the full list once, which also orders the tail the output needs.
template <int R, int HPL>
__device__ bool merge_misses(uint64_t* keys, uint32_t Km, const uint64_t* tail,
                             uint32_t m, uint64_t* ms, bool force_final) {
  const uint32_t lane = lane_id();
  uint64_t mv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t idx = lane * R + r;
    mv[r] = idx < m ? tail[idx] : kPadKey;
  }
  reg_bitonic_sort<R>(mv);
  wave_sync();  // ms overlays the todo list
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t idx = lane * R + r;
    if (idx < m) ms[idx] = mv[r];
  }
  wave_sync();
  constexpr int kHpl = HPL;  // Km <= 64 * HPL
  uint64_t hx[kHpl];
  uint32_t hp[kHpl];
  bool unexp = false;
#pragma unroll
  for (int s = 0; s < kHpl; ++s) {
    const uint32_t i = lane + 64u * s;
    hx[s] = 0;
    hp[s] = 0xFFFFFFFFu;
    if (i < Km) {
      hx[s] = keys[i];
      hp[s] = i + count_less(ms, m, hx[s]);
      if (hp[s] < Km && (hx[s] & 1ull) == 0) unexp = true;
    }
  }
  uint32_t mpos[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t idx = lane * R + r;
    mpos[r] = 0xFFFFFFFFu;
    if (idx < m) {
      mpos[r] = idx + count_less_eq(keys, Km, mv[r]);
      if (mpos[r] < Km) unexp = true;
    }
  }
  if (force_final || __ballot(unexp) == 0) return true;
  wave_sync();
#pragma unroll
  for (int s = 0; s < kHpl; ++s)
    if (hp[s] < Km) keys[hp[s]] = hx[s];
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (mpos[r] < Km) keys[mpos[r]] = mv[r];
  wave_sync();
  return false;
}

// Merge by ranks, for any number of entering misses (keys tail[0, m)).  The
// misses below the current Km-th key T enter (keys are unique per id); they are compacted into
// ebuf, each one's position is #(entering < e) + #(kept < e) (a scan of ebuf
// plus a binary search, one lane per entering key), and they are placed into
// a staging copy of the top Km held in the dead tail region (it must hold Km
// keys).  The kept keys then fill the free slots in order: free slot p takes
// old key p - #(occupied slots before p).  Same list as merge_misses.
// Returns 1 when this iteration is the last one (nothing enters and no
// unexpanded key is in the top Km, or force_final), else 0.
template <int CH, int HPL>
__device__ int merge_ranked(uint64_t* keys, uint32_t Km, uint64_t* stage,
                            const uint64_t* tail, uint32_t m, uint64_t* ebuf,
                            bool force_final) {
  const uint32_t lane = lane_id();
  const uint64_t T = keys[Km - 1];
  uint64_t me[CH];
  uint32_t slot[CH];
  bool in_e[CH];
  uint32_t ne = 0;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    me[c] = kPadKey;
    in_e[c] = false;
    slot[c] = 0;
    if (64u * c < m) {
      const uint32_t idx = lane + 64u * c;
      if (idx < m) me[c] = tail[idx];
      in_e[c] = me[c] < T;
      const uint64_t em = __ballot(in_e[c]);
      slot[c] = ne + __popcll(em & lanemask_lt());
      ne += __popcll(em);
    }
  }
  bool unexp = false;
  uint64_t k0 = kPadKey;  // top-list key of this lane (slice 0)
#pragma unroll
  for (int s = 0; s < HPL; ++s) {
    const uint32_t i = lane + 64u * s;
    if (i < Km) {
      const uint64_t k = keys[i];
      if (s == 0) k0 = k;
      if ((k & 1ull) == 0) unexp = true;
    }
  }
  if (ne == 0) return (force_final || __ballot(unexp) == 0) ? 1 : 0;
  if (force_final) return 1;
  if constexpr (HPL == 1 && CH == 1) {
    // Few entering keys (the common case at narrow beams): one uniform
    // pass per entering key e (taken from its lane) finds its final
    // position #(kept < e) + #(entering < e) from two ballots; lane p then
    // takes the entering key placed at p, or else old key p - #(entering
    // keys placed before p).  Same list as the general path below.
    constexpr uint32_t kFewEntering = 8;
    if (ne <= kFewEntering) {
      uint64_t em = __ballot(in_e[0]);
      uint32_t before = 0;
      uint64_t sel = 0;
      bool hit = false;
      while (em) {
        const int l = __ffsll((long long)em) - 1;
        em &= em - 1;
        const uint64_t e =
            (uint64_t(__builtin_amdgcn_readlane(uint32_t(me[0] >> 32), l))
             << 32) |
            uint32_t(__builtin_amdgcn_readlane(uint32_t(me[0]), l));
        const uint32_t pos = __popcll(__ballot(lane < Km && k0 < e)) +
                             __popcll(__ballot(in_e[0] && me[0] < e));
        before += pos < lane ? 1u : 0u;
        if (pos == lane) {
          sel = e;
          hit = true;
        }
      }
      // every lane reads its old key before any lane writes (the write
      // depends on the read)
      const uint64_t v = hit ? sel : (lane < Km ? keys[lane - before] : 0);
      if (lane < Km) keys[lane] = v;
      wave_sync();
      return 0;
    }
  }
  wave_sync();  // ebuf overlays the todo list
#pragma unroll
  for (int c = 0; c < CH; ++c)
    if (in_e[c]) ebuf[slot[c]] = me[c];
#pragma unroll
  for (int s = 0; s < HPL; ++s) {
    const uint32_t p = lane + 64u * s;
    if (p < Km) stage[p] = kPadKey;
  }
  wave_sync();
  for (uint32_t e0 = 0; e0 < ne; e0 += kWave) {
    const bool own = e0 + lane < ne;
    const uint64_t e = own ? ebuf[e0 + lane] : kPadKey;
    uint32_t r = 0;
#pragma unroll 4
    for (uint32_t j = 0; j < ne; ++j) r += ebuf[j] < e ? 1u : 0u;
    if (own) {
      r += count_less(keys, Km, e);
      if (r < Km) stage[r] = e;
    }
  }
  wave_sync();
  uint64_t v[HPL];
  uint32_t occ_before = 0;
#pragma unroll
  for (int s = 0; s < HPL; ++s) {
    const uint32_t p = lane + 64u * s;
    v[s] = p < Km ? stage[p] : kPadKey;
    const bool occ = p < Km && v[s] != kPadKey;
    const uint64_t om = __ballot(occ);
    const uint32_t before = occ_before + __popcll(om & lanemask_lt());
    if (p < Km && !occ) v[s] = keys[p - before];
    occ_before += __popcll(om);
  }
  wave_sync();
#pragma unroll
  for (int s = 0; s < HPL; ++s) {
    const uint32_t p = lane + 64u * s;
    if (p < Km) keys[p] = v[s];
  }
  wave_sync();
  return 0;
}

// HPL: top-list slices of 64 keys per lane (Km <= 64*HPL), CH: candidate
// slices (Kp*Kd <= 64*CH).  Compile-time slice counts keep the merge free of
// dead slices; larger lists fall back to a full sort per iteration.
//
// Waves per query: the workgroup is blockDim.x / 64 waves.  Wave 0 runs the
// search; any further waves only evaluate distances: at each evaluation
// wave 0 publishes (command, query, count) in LDS and meets them at a
// workgroup barrier, every wave evaluates its share of the candidates, and
// a second barrier hands the keys back.  Between those barriers wave 0
// orders its own LDS traffic with wave_sync, so the helpers, which wait at
// the next barrier, never see its other phases.  Command 0 ends the helpers
// (after wave 0's last query), so every wave passes the same barriers.
//
// kLean: a variant for the shapes the fast paths always cover (the host
// picks it when Km <= 64*HPL, Kp*Kd <= 64*CH, the tail can stage Km keys,
// max_iters > 0 and one wave per query): only the rank merge and the
// register entry sort are compiled, not the fallbacks nor the evaluation
// hand-off between waves, which leaves the compiler fewer live values in
// the iteration loop.  Same lists.
template <class Eval, int HPL, int CH, bool kLean = false>
__global__ __launch_bounds__(128)
__attribute__((amdgpu_waves_per_eu(kLean ? Eval::kLeanMinWaves
                                         : Eval::kMinWaves, 8))) void search_kernel(
    Eval ev, const float* __restrict__ queries, uint32_t n_query,
    uint32_t n_base,
    const uint32_t* __restrict__ graph, uint32_t gdeg, uint32_t Km,
    uint32_t Kp, uint32_t Kd, uint32_t topk, uint32_t reset_iter,
    uint32_t entry_seed, uint32_t max_iters, uint32_t tsize, uint32_t lp2,
    uint32_t kd_shift, uint32_t* __restrict__ out,
    unsigned long long* __restrict__ stats, uint32_t* __restrict__ queue,
    const uint32_t* __restrict__ order) {
  extern __shared__ __align__(16) uint8_t smem[];
  const uint32_t kpkd = Kp * Kd;
  ev.qs = reinterpret_cast<bf16x8*>(smem);
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem + Eval::kQueryLdsBytes);
  uint64_t* keys = reinterpret_cast<uint64_t*>(
      smem + Eval::kQueryLdsBytes + ((size_t(tsize) * 4 + 15) & ~size_t(15)));
  uint32_t* mscratch = reinterpret_cast<uint32_t*>(keys + lp2);
  uint32_t* todo = mscratch + kpkd;
  uint32_t* explore = todo + kpkd;
  // the evaluation hand-off words: command, query, candidate count
  volatile uint32_t* hand = explore + Kp;
  // sorted misses of the merge step: mscratch + todo (2*kpkd words)
  uint64_t* ms = reinterpret_cast<uint64_t*>(mscratch);
  const uint32_t lane = lane_id();
  // the lean variant runs one wave per workgroup
  const uint32_t wave = kLean ? 0u : threadIdx.x >> 6;
  const uint32_t nw = kLean ? 1u : blockDim.x >> 6;
  uint64_t* tail = keys + Km;
  if (wave != 0) {
    uint32_t cur_q = 0xFFFFFFFFu;
    while (true) {
      __syncthreads();  // wave 0 has published a command
      const uint32_t cmd = hand[0];
      if (cmd == 0) break;
      const uint32_t qq = hand[1], nt = hand[2];
      if (qq != cur_q) {
        ev.load_query_regs(queries, qq);
        cur_q = qq;
      }
      ev.eval(todo, nt, tail, wave, nw);
      __syncthreads();  // keys handed back
    }
    return;
  }
  // One query per workgroup, or (queue != nullptr) a persistent grid whose
  // waves take query ids from a device counter until it runs past n_query.
  // With a locality order, the queries are taken in that order (nodes
  // grouped by nearest pivot row, so that concurrent searches gather
  // overlapping rows), split into kQueueSegs contiguous segments with a
  // counter each.  A wave starts on segment blockIdx % kQueueSegs, which is
  // the XCD its workgroup is dispatched to (round-robin), and moves to the
  // next segment once its own is drained; segments only drain, so a wave
  // passes each empty segment once and the loop ends.
  const uint32_t seg_len = (n_query + kQueueSegs - 1) / kQueueSegs;
  uint32_t seg = blockIdx.x % kQueueSegs, segs_left = kQueueSegs;
  auto next_query = [&]() -> uint32_t {
    uint32_t v = n_query;
    if (lane == 0) {
      if (order == nullptr) {
        v = atomicAdd(queue, 1u);
      } else {
        while (segs_left > 0) {
          const uint32_t lo = seg * seg_len;
          const uint32_t len =
              lo >= n_query ? 0u : (n_query - lo < seg_len ? n_query - lo : seg_len);
          if (len > 0) {
            const uint32_t i = atomicAdd(&queue[seg], 1u);
            if (i < len) {
              v = order[lo + i];
              break;
            }
          }
          seg = seg + 1u == kQueueSegs ? 0u : seg + 1u;
          --segs_left;
        }
      }
    }
    return __shfl(v, 0, 64);
  };
  uint32_t q = queue ? next_query() : blockIdx.x;
  // distances of todo[0, ntodo): all waves of the workgroup take part
  auto eval_all = [&](uint32_t ntodo) {
    if (nw > 1) {
      if (lane == 0) {
        hand[1] = q;
        hand[2] = ntodo;
        hand[0] = 1u;
      }
      __syncthreads();
      ev.eval(todo, ntodo, tail, 0, nw);
      __syncthreads();
    } else {
      wave_sync();
      ev.eval(todo, ntodo, tail);
      wave_sync();
    }
  };
  while (q < n_query) {
  const bool fast_merge = kLean || (Km <= 64u * HPL && kpkd <= 64u * CH);
  // the rank merge stages the new top list in the tail region
  const bool rank_merge = kLean || (fast_merge && lp2 - Km >= Km);
  ev.load_query(queries, q);
  // initial list: [0, Km) = (FLT_MAX, tomb); [Km, Km+kpkd) = entry points
  for (uint32_t i = lane; i < lp2; i += kWave)
    keys[i] = i < Km ? search_key(FLT_MAX, kTomb) : kPadKey;
  for (uint32_t t = lane; t < kpkd; t += kWave)
    todo[t] = uint32_t((uint64_t(entry_seed) + t) % n_base);
  eval_all(kpkd);
  unsigned long long ndist = kpkd;
  // The list after init is the first Km keys of sort(entry keys, Km tomb
  // keys): the sorted entry keys below the tomb key, then tomb keys.  The
  // tail is rewritten by the first iteration before anything reads it, so
  // only the entry keys need sorting (in registers).  With max_iters == 0
  // the initial list is the output, which needs the full sort.
  constexpr int RS = CH <= 1 ? 1 : CH <= 2 ? 2 : 4;
  if (kLean || (max_iters > 0 && kpkd <= 64u * RS && Km <= 64u * RS)) {
    const uint64_t tomb_key = search_key(FLT_MAX, kTomb);
    uint64_t v[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const uint32_t e = lane * RS + r;
      v[r] = e < kpkd ? tail[e] : kPadKey;
    }
    reg_bitonic_sort<RS>(v);
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const uint32_t e = lane * RS + r;
      if (e < Km) keys[e] = v[r] < tomb_key ? v[r] : tomb_key;
    }
    wave_sync();
  } else if constexpr (!kLean) {
    bitonic_sort<uint64_t, true>(keys, lp2);
  }
  auto reset_table = [&]() {
    // 16-byte stores (the table starts 16-byte aligned), then the rest
    for (uint32_t i = lane; i < (tsize >> 2); i += kWave)
      reinterpret_cast<uint4*>(tab)[i] = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t i = (tsize & ~3u) + lane; i < tsize; i += kWave) tab[i] = 0u;
    wave_sync();
    for (uint32_t i = lane; i < Km; i += kWave)
      visit_insert(tab, tsize, key_raw_id(keys[i]) & kMaskedTomb);
    wave_sync();
  };
  reset_table();
  // Exact evaluators keep the reference's visited set (reset every
  // reset_iter iterations).  The screened evaluator treats the table as a
  // cache: it is cleared and re-seeded with the top-Km ids only when the
  // next iteration could push it past 3/4 full.  The top-Km ids are always
  // in it, so no id enters the list twice; a dropped id can only be
  // evaluated again.
  const uint32_t cache_limit = tsize - tsize / 4;
  uint32_t cache_count = Km;
  uint32_t iter = 0, total_iter = 0;
  bool capped = false;
  while (true) {
    // collect_top_p_normal (Gsearcher.cuh:75-106): first Kp unexpanded
    uint32_t found = 0;
    for (uint32_t c0 = 0; c0 < Km && found < Kp; c0 += kWave) {
      const uint32_t i = c0 + lane;
      uint64_t k = 0;
      bool un = false;
      if (i < Km) {
        k = keys[i];
        un = (k & 1ull) == 0;
      }
      const uint64_t m = __ballot(un);
      const uint32_t rank = __popcll(m & lanemask_lt());
      if (un && found + rank < Kp) {
        explore[found + rank] = key_raw_id(k);
        keys[i] = k | 1ull;
      }
      const uint32_t pc = __popcll(m);
      found += pc < Kp - found ? pc : Kp - found;
    }
    if (found == 0) break;
    if (total_iter >= max_iters) {
      capped = true;
      break;
    }
    for (uint32_t e = found + lane; e < Kp; e += kWave) explore[e] = kTomb;
    wave_sync();
    if (!Eval::kExactVisited && cache_count + kpkd > cache_limit) {
      reset_table();
      cache_count = Km;
    }
    // neighbour ids of the expanded nodes (first Kd of each graph row),
    // all slices in flight together, then the visited test-and-set slice
    // by slice from registers; misses go to the todo list, hits get FLT_MAX
    uint32_t nid[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint32_t t = lane + 64u * c;
      nid[c] = kTomb;
      if (t < kpkd) {
        const uint32_t e = kd_shift < 32u ? (t >> kd_shift) : t / Kd;
        const uint32_t j = t - e * Kd;
        const uint32_t eid = explore[e];
        if (eid != kTomb) nid[c] = graph[size_t(eid) * gdeg + j];
      }
    }
    // Test-and-set of all slices in lockstep probe rounds, so that their
    // LDS atomics are in flight together.  Slice c's CAS is issued before
    // slice c+1's in every round and a duplicate id walks the same probe
    // sequence, so an id in two slices is a miss in the earlier one, as
    // with one slice after the other.
    bool vmiss[CH], vact[CH];
    uint32_t vslot[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint32_t t = lane + 64u * c;
      vmiss[c] = false;
      vact[c] = t < kpkd && nid[c] != kTomb;
      vslot[c] = visit_hash(nid[c], tsize);
    }
    while (true) {
      bool any = false;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        if (vact[c]) {
          const uint32_t old = atomicCAS(&tab[vslot[c]], 0u, nid[c] + 1u);
          if (old == 0u) {
            vmiss[c] = true;
            vact[c] = false;
          } else if (old == nid[c] + 1u) {
            vact[c] = false;
          } else {
            vslot[c] = vslot[c] + 1u == tsize ? 0u : vslot[c] + 1u;
          }
        }
        any = any || vact[c];
      }
      if (!any) break;
    }
    // The misses' ids are compacted into todo and their keys go to
    // tail[0, ntodo); the hits' (FLT_MAX) keys fill the rest of tail[0,
    // kpkd): the tail holds the same keys as slot by slot, which is all the
    // final sort sees.
    uint32_t ntodo = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const bool miss = vmiss[c];
      // (a hit's FLT_MAX key is written only on the last iteration, below)
      const uint64_t m = __ballot(miss);
      if (miss) todo[ntodo + __popcll(m & lanemask_lt())] = nid[c];
      ntodo += __popcll(m);
    }
    // candidate slots past 64*CH (Kp*Kd up to 4096), one slice at a time;
    // their hits' keys go to the end of tail[0, kpkd)
    uint32_t nwh = 0;
    for (uint32_t c0 = 64u * CH; !kLean && c0 < kpkd; c0 += kWave) {
      const uint32_t t = c0 + lane;
      bool miss = false;
      uint32_t id = kTomb;
      if (t < kpkd) {
        const uint32_t e = kd_shift < 32u ? (t >> kd_shift) : t / Kd;
        const uint32_t j = t - e * Kd;
        const uint32_t eid = explore[e];
        id = eid == kTomb ? kTomb : graph[size_t(eid) * gdeg + j];
        if (id != kTomb) miss = visit_insert(tab, tsize, id);
      }
      const uint64_t m = __ballot(miss);
      if (miss) todo[ntodo + __popcll(m & lanemask_lt())] = id;
      ntodo += __popcll(m);
      const bool hit = t < kpkd && !miss;
      const uint64_t hm = __ballot(hit);
      if (hit)
        tail[kpkd - 1u - (nwh + __popcll(hm & lanemask_lt()))] =
            search_key(FLT_MAX, id);
      nwh += __popcll(hm);
    }
    eval_all(ntodo);
    ndist += ntodo;
    cache_count += ntodo;
    const bool last_allowed = total_iter + 1 >= max_iters;
    bool final_sort = true;
    if (kLean || rank_merge) {
      final_sort = merge_ranked<CH, HPL>(keys, Km, tail, tail, ntodo, ms,
                                         last_allowed) == 1;
    } else if (!kLean && fast_merge) {
      if (ntodo <= 64)
        final_sort = merge_misses<1, HPL>(keys, Km, tail, ntodo, ms,
                                        last_allowed);
      else if (CH <= 2 || ntodo <= 128)
        final_sort = merge_misses<(CH < 2 ? CH : 2), HPL>(keys, Km, tail,
                                                      ntodo, ms, last_allowed);
      else  // the register bitonic sort needs a power-of-two width
        final_sort = merge_misses<(CH <= 4 ? 4 : 8), HPL>(
            keys, Km, tail, ntodo, ms, last_allowed);
    }
    if (final_sort) {
      // The output's tail is the sorted union of the list and this last
      // iteration's candidates, its visited hits at FLT_MAX: their keys
      // are written here, after the misses' (tail[0, ntodo)).
      uint32_t hpos = ntodo;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const uint32_t t = lane + 64u * c;
        const bool hit = t < kpkd && !vmiss[c];
        const uint64_t hm = __ballot(hit);
        if (hit)
          tail[hpos + __popcll(hm & lanemask_lt())] =
              search_key(FLT_MAX, nid[c]);
        hpos += __popcll(hm);
      }
      wave_sync();
      sort_list<true>(keys, lp2);
    }
    ++total_iter;
    if (Eval::kExactVisited && ++iter == reset_iter) {
      reset_table();
      iter = 0;
    }
  }
  for (uint32_t i = lane; i < topk; i += kWave)
    out[size_t(q) * topk + i] = key_raw_id(keys[i]) & kMaskedTomb;
  if (lane == 0) {
    stat_add(stats, ST_SEARCH_DIST, ndist);
    stat_add(stats, ST_SEARCH_ITERS, total_iter);
    if (capped) stat_add(stats, ST_SEARCH_CAPPED, 1);
  }
  if (queue == nullptr) break;
  wave_sync();
  q = next_query();
  }  // query loop
  if (nw > 1) {
    if (lane == 0) hand[0] = 0u;
    __syncthreads();  // the helpers stop
  }
}


template <class Eval>
static hipError_t launch_search_eval(
    const Eval& ev, const float* queries, uint32_t n_query, uint32_t n_base,
    const uint32_t* graph, uint32_t gdeg, uint32_t Km, uint32_t Kp,
    uint32_t Kd, uint32_t topk, uint32_t reset_iter, uint32_t entry_seed,
    uint32_t max_iters, uint32_t table_size, uint32_t* out,
    unsigned long long* stats, uint32_t* queue, const uint32_t* order,
    uint32_t waves, hipStream_t s) {
  const uint32_t block = 64u * (waves == 2 ? 2u : 1u);
  uint32_t lp2 = 64;
  while (lp2 < Km + Kp * Kd) lp2 <<= 1;
  const size_t smem = search_smem_bytes(table_size, lp2, Kp * Kd, Kp) +
                      Eval::kQueryLdsBytes;
  uint32_t kd_shift = 0xFFu;
  for (uint32_t sh = 0; sh < 31; ++sh)
    if ((1u << sh) == Kd) kd_shift = sh;
  const uint32_t kpkd = Kp * Kd;
  // top-list / candidate slices per lane: 1/1 up to 64 entries each (the
  // narrow screened beams; fewer live registers, no spills), 2/2 up to 128,
  // 3/3 up to 192, else 4/4 (lists beyond 256 sort fully every iteration)
  const int slices = (Km <= 64 && kpkd <= 64) ? 1
                     : (Km <= 128 && kpkd <= 128) ? 2
                     : (Km <= 192 && kpkd <= 192) ? 3 : 4;
  // the lean one-slice variant where its fast paths always apply
  const bool lean =
      slices == 1 && max_iters > 0 && lp2 - Km >= Km && block == 64u;
  const void* fn =
      lean ? reinterpret_cast<const void*>(&search_kernel<Eval, 1, 1, true>)
      : slices == 1 ? reinterpret_cast<const void*>(&search_kernel<Eval, 1, 1>)
      : slices == 2 ? reinterpret_cast<const void*>(&search_kernel<Eval, 2, 2>)
      : slices == 3 ? reinterpret_cast<const void*>(&search_kernel<Eval, 3, 3>)
                    : reinterpret_cast<const void*>(&search_kernel<Eval, 4, 4>);
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(smem));
  uint32_t grid = n_query;
  if (queue) {
    // persistent grid: as many one-wave workgroups as can be resident
    int dev = 0, per_cu = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block,
                                                       smem);
    if (per_cu < 1) per_cu = 1;
    if (cus < 1) cus = 1;
    const uint64_t resident = uint64_t(per_cu) * uint64_t(cus);
    grid = uint32_t(resident < n_query ? resident : n_query);
    hipError_t e = launch_fill_u32(queue, kQueueSegs, 0u, s);
    if (e != hipSuccess) return e;
  }
#define PG_SEARCH_LAUNCH(H, C, LEAN)                                        \
  search_kernel<Eval, H, C, LEAN><<<grid, block, smem, s>>>(                \
      ev, queries, n_query, n_base, graph, gdeg, Km, Kp, Kd, topk, reset_iter, \
      entry_seed, max_iters, table_size, lp2, kd_shift, out, stats, queue,  \
      queue ? order : nullptr)
  if (lean) PG_SEARCH_LAUNCH(1, 1, true);
  else if (slices == 1) PG_SEARCH_LAUNCH(1, 1, false);
  else if (slices == 2) PG_SEARCH_LAUNCH(2, 2, false);
  else if (slices == 3) PG_SEARCH_LAUNCH(3, 3, false);
  else PG_SEARCH_LAUNCH(4, 4, false);
#undef PG_SEARCH_LAUNCH
  return hipGetLastError();
}

template <int TQ, int U>
static hipError_t launch_f32(const float* base, uint32_t dim,
                             const float* queries, uint32_t n_query,
                             uint32_t n_base, const uint32_t* graph,
                             uint32_t gdeg, uint32_t Km, uint32_t Kp,
                             uint32_t Kd, uint32_t topk, uint32_t reset_iter,
                             uint32_t entry_seed, uint32_t max_iters,
                             uint32_t table_size, uint32_t* out,
                             unsigned long long* stats, uint32_t* queue,
                             const uint32_t* order, uint32_t waves,
                             hipStream_t s) {
  EvalF32<TQ, U> ev;
  ev.base = base;
  ev.dim = dim;
  return launch_search_eval(ev, queries, n_query, n_base, graph, gdeg, Km, Kp,
                            Kd, topk, reset_iter, entry_seed, max_iters,
                            table_size, out, stats, queue, order, waves, s);
}

template <int NCH, int G>
static hipError_t launch_bf16(const void* base16, uint32_t dim, uint32_t dpad,
                              const float* queries, uint32_t n_query,
                              uint32_t n_base, const uint32_t* graph,
                              uint32_t gdeg, uint32_t Km, uint32_t Kp,
                              uint32_t Kd, uint32_t topk, uint32_t reset_iter,
                              uint32_t entry_seed, uint32_t max_iters,
                              uint32_t table_size, uint32_t* out,
                              unsigned long long* stats, uint32_t* queue,
                              const uint32_t* order, uint32_t waves,
                              hipStream_t s) {
  EvalBF16<NCH, G> ev;
  ev.base16 = reinterpret_cast<const __bf16*>(base16);
  ev.dim = dim;
  ev.dpad = dpad;
  ev.nch = dpad / 32;
  return launch_search_eval(ev, queries, n_query, n_base, graph, gdeg, Km, Kp,
                            Kd, topk, reset_iter, entry_seed, max_iters,
                            table_size, out, stats, queue, order, waves, s);
}

template <int NCH, int G>
static hipError_t launch_fp8(const void* base8, const float* scale8,
                             uint32_t dim, uint32_t dpad, const float* queries,
                             uint32_t n_query, uint32_t n_base,
                             const uint32_t* graph, uint32_t gdeg, uint32_t Km,
                             uint32_t Kp, uint32_t Kd, uint32_t topk,
                             uint32_t reset_iter, uint32_t entry_seed,
                             uint32_t max_iters, uint32_t table_size,
                             uint32_t* out, unsigned long long* stats,
                             uint32_t* queue, const uint32_t* order,
                             uint32_t waves, hipStream_t s) {
  EvalFP8<NCH, G> ev;
  ev.base8 = reinterpret_cast<const uint8_t*>(base8);
  ev.scale8 = scale8;
  ev.dim = dim;
  ev.dpad = dpad;
  ev.nch = dpad / 32;
  return launch_search_eval(ev, queries, n_query, n_base, graph, gdeg, Km, Kp,
                            Kd, topk, reset_iter, entry_seed, max_iters,
                            table_size, out, stats, queue, order, waves, s);
}

// Launch parameters shared by every evaluator family.
#define PG_SEARCH_PARAMS                                                    \
  const float *queries, uint32_t n_query, uint32_t n_base,                  \
      const uint32_t *graph, uint32_t gdeg, uint32_t Km, uint32_t Kp,       \
      uint32_t Kd, uint32_t topk, uint32_t reset_iter, uint32_t entry_seed, \
      uint32_t max_iters, uint32_t table_size, uint32_t *out,               \
      unsigned long long *stats, uint32_t *queue, const uint32_t *order,    \
      uint32_t waves, hipStream_t s
#define PG_SEARCH_ARGS                                                      \
  queries, n_query, n_base, graph, gdeg, Km, Kp, Kd, topk, reset_iter,      \
      entry_seed, max_iters, table_size, out, stats, queue, order, waves, s

// Exact FP32 search, query slices of TQ registers (32*TQ >= dim) and
// gather_batch u candidate pairs in flight (4, 8 or 16):
// pg_search_f32_<TQ>.hip.
hipError_t launch_search_f32_4(const float* base, uint32_t dim, uint32_t u,
                               PG_SEARCH_PARAMS);
hipError_t launch_search_f32_8(const float* base, uint32_t dim, uint32_t u,
                               PG_SEARCH_PARAMS);
hipError_t launch_search_f32_16(const float* base, uint32_t dim, uint32_t u,
                                PG_SEARCH_PARAMS);
hipError_t launch_search_f32_32(const float* base, uint32_t dim, uint32_t u,
                                PG_SEARCH_PARAMS);
// BF16- and FP8-screened searches (pg_search_bf16.hip, pg_search_fp8.hip):
// nch = dpad / 32 chunks (<= 16), two = two 16-candidate groups in flight.
hipError_t launch_search_bf16(const void* base16, uint32_t dim, uint32_t dpad,
                              bool two, PG_SEARCH_PARAMS);
hipError_t launch_search_fp8(const void* base8, const float* scale8,
                             uint32_t dim, uint32_t dpad, bool two,
                             PG_SEARCH_PARAMS);

}  // namespace pg
