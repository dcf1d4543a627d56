// cpu_ref — CPU restatement of ParaGraph's GPU index-build stage.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker.  The product path (paragraph_amd)
// never links, loads or calls it.
//
// Every function restates the semantics of one reference routine (file:line
// relative to the reference repository).  Where the reference leaves a result
// to hardware scheduling, the canonical rules below fix it; the GPU kernels
// follow the same rules, so GPU == oracle bit-exactly on neighbour ids.
//
//   C1  top-1 match: the lowest query id wins a contested base row
//       (reference: atomicCAS winner, include/Gbuilder.cuh:122).
//   C2  reverse lists: the sources of a target are kept in ascending source
//       id, truncated to the cap (reference: atomicAdd order + racy clamp,
//       include/Gbuilder.cuh:1020-1038).
//   C3  search list order: (distance, id & 0x7FFFFFFF, expanded-bit)
//       ascending (reference: warp_sort on shared memory, whose tie order is
//       unspecified, include/GBitonicSort.cuh:237-244).
//   C5  ids >= n_base that are not the tomb / masked tomb are treated as tomb
//       (the reference would read out of bounds).
//
// Arithmetic: distances are FP32 with explicit fused multiply-add, matching
// nvcc's default contraction of `sum += x*y`:
//   chain_dist  = -(sequential FMA chain over i = 0..dim-1)   Gbuilder.cuh:785-797
//   search_dist = -(32 lane-strided FMA chains + 16/8/4/2/1 tree)
//                                                   Gsearcher.cuh:958-976,1023-1035
//
// Pinning: tests/test_oracle.py checks each function against hand-derived
// tiny cases, brute force (exhaustive search on complete graphs, RNG on
// <=12 points), FP64 error bounds on the distances, and invariants.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr uint32_t kTomb = 0xFFFFFFFFu;
constexpr uint32_t kMaskedTomb = 0x7FFFFFFFu;
constexpr uint32_t kMsb = 0x80000000u;

inline float chain_dist(const float* x, const float* y, uint32_t dim) {
  // Gbuilder.cuh:785-797 — sum starts at 0, one FMA per element, negated.
  float s = 0.0f;
  for (uint32_t i = 0; i < dim; ++i) s = std::fmaf(x[i], y[i], s);
  return -s;
}

inline float search_dist(const float* q, const float* x, uint32_t dim) {
  // Gsearcher.cuh:1023-1035 — lane l accumulates i = l, l+32, ... then
  // __shfl_down_sync over offsets 16, 8, 4, 2, 1; lane 0 holds the sum.
  float p[32];
  for (uint32_t l = 0; l < 32; ++l) {
    float s = 0.0f;
    for (uint32_t i = l; i < dim; i += 32) s = std::fmaf(q[i], x[i], s);
    p[l] = s;
  }
  for (uint32_t o = 16; o > 0; o >>= 1)
    for (uint32_t l = 0; l < o; ++l) p[l] = p[l] + p[l + o];
  return -p[0];
}

inline bool is_stop_id(uint32_t id, uint32_t n) {
  // tomb, masked tomb (INT_MAX) and out-of-range ids (C5)
  return id == kTomb || id == kMaskedTomb || id >= n;
}

inline uint32_t ordered_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u & kMsb) ? ~u : (u | kMsb);
}

inline uint32_t rotl1(uint32_t x) { return (x << 1) | (x >> 31); }

inline uint64_t search_key(float d, uint32_t raw_id) {
  // C3: (distance, id & 0x7FFFFFFF, expanded bit).
  return (uint64_t(ordered_bits(d)) << 32) | rotl1(raw_id);
}

struct Cand {
  float d;
  uint32_t id;
};

// Stable ascending sort by distance (Gbuilder.cuh:890-902 bubble_sort swaps
// only when K[j] > K[j+1], so it is exactly a stable sort).
inline void stable_by_dist(std::vector<Cand>& v) {
  std::stable_sort(v.begin(), v.end(),
                   [](const Cand& a, const Cand& b) { return a.d < b.d; });
}

// Strict RNG occlusion over an already sorted list (Gbuilder.cuh:1150-1198 and
// :1500-1553).  `kept` holds the list entries kept so far; candidate i is
// occluded iff some kept k (checked in insertion order, early exit) has
// chain_dist(c_i, k) < d_i.  Returns the number of pair distances evaluated.
inline uint64_t occluded_by(const float* base, uint32_t dim, const Cand& c,
                            const std::vector<uint32_t>& kept, bool* occl) {
  uint64_t pairs = 0;
  *occl = false;
  const float* xc = base + size_t(c.id) * dim;
  for (uint32_t k : kept) {
    ++pairs;
    float pd = chain_dist(xc, base + size_t(k) * dim, dim);
    if (pd < c.d) {
      *occl = true;
      break;
    }
  }
  return pairs;
}

}  // namespace

extern "C" {

// Counters shared with the GPU path (same meaning, see pg_stats in
// include/paragraph.h).
struct orc_stats {
  uint64_t prune_dist;     // node->candidate distances in prune (D_valid)
  uint64_t prune_pairs;    // occlusion pair distances in prune (P)
  uint64_t merge_dist;     // node->candidate distances in merge (C)
  uint64_t merge_pairs;    // occlusion pair distances in merge
  uint64_t search_dist;    // hashtable-miss distance evaluations (E)
  uint64_t search_iters;   // search iterations (I)
  uint64_t search_capped;  // queries stopped by the iteration watchdog
  uint64_t rev_overflow;   // targets whose reverse list exceeded the cap
};

int orc_version() { return 1; }

float orc_chain_dist(const float* x, const float* y, uint32_t dim) {
  return chain_dist(x, y, dim);
}
float orc_search_dist(const float* q, const float* x, uint32_t dim) {
  return search_dist(q, x, dim);
}

// K0 match_top1_kernel_v0 (Gbuilder.cuh:104-135) under C1: winner[b] is the
// lowest query q with gt[q][0] == b.
void orc_match_winner(const uint32_t* gt, uint64_t n_query, uint32_t gt_k,
                      uint32_t n_base, uint32_t* winner) {
  for (uint32_t b = 0; b < n_base; ++b) winner[b] = kTomb;
  for (uint64_t q = 0; q < n_query; ++q) {
    uint32_t b = gt[q * gt_k];
    if (b < n_base && winner[b] == kTomb) winner[b] = uint32_t(q);
  }
}

// The 128-wide top-1 row the winner writes (Gbuilder.cuh:121-132 with
// match_real_degree = gt_k - 1): slots [0, gt_k-2) = gt[w][1 .. gt_k-2],
// the last two slots tomb; a base nobody matched keeps the all-tomb row
// the host initialised (tests/test_t2iend.cu:866).
void orc_top1_rows(const uint32_t* gt, uint32_t gt_k, const uint32_t* winner,
                   uint32_t n_base, uint32_t* rows) {
  for (uint32_t b = 0; b < n_base; ++b) {
    uint32_t* r = rows + size_t(b) * gt_k;
    uint32_t w = winner[b];
    for (uint32_t j = 0; j < gt_k; ++j) r[j] = kTomb;
    if (w == kTomb) continue;
    for (uint32_t j = 0; j + 2 < gt_k; ++j) r[j] = gt[size_t(w) * gt_k + j + 1];
  }
}

// K2 fusion_prune_reverse_kernel_v0 (Gbuilder.cuh:1105-1244), pruned part:
// per node, distances to the D row candidates (tomb / self / INT_MAX ->
// FLT_MAX), stable sort, node skipped if the first entry is tomb/INT_MAX,
// then strict RNG keeping at most `mp` (the loop stops at the first
// tomb/INT_MAX).  The non-strict "slight" fill (:1201-1239) only visits
// FLT_MAX entries, all of which it skips, so it adds nothing and is omitted.
// Reverse edges are derived from the pruned lists by orc_reverse_csr.
void orc_prune(const float* base, uint32_t n_base, uint32_t dim,
               const uint32_t* rows, uint32_t D, uint32_t mp, uint32_t* pruned,
               uint32_t* pcnt, orc_stats* st) {
  uint64_t tot_dist = 0, tot_pairs = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tot_dist, tot_pairs)
  for (int64_t bi = 0; bi < int64_t(n_base); ++bi) {
    uint32_t b = uint32_t(bi);
    const uint32_t* row = rows + size_t(b) * D;
    const float* xb = base + size_t(b) * dim;
    std::vector<Cand> c(D);
    for (uint32_t j = 0; j < D; ++j) {
      uint32_t id = row[j];
      c[j].id = id;
      if (id == kTomb || id == b || id == kMaskedTomb || id >= n_base) {
        c[j].d = FLT_MAX;
      } else {
        c[j].d = chain_dist(xb, base + size_t(id) * dim, dim);
        ++tot_dist;
      }
    }
    stable_by_dist(c);
    uint32_t* out = pruned + size_t(b) * mp;
    pcnt[b] = 0;
    if (is_stop_id(c[0].id, n_base)) continue;
    std::vector<uint32_t> kept;
    kept.reserve(mp);
    kept.push_back(c[0].id);
    for (uint32_t i = 1; i < D && kept.size() < mp; ++i) {
      if (is_stop_id(c[i].id, n_base)) break;
      bool occl;
      tot_pairs += occluded_by(base, dim, c[i], kept, &occl);
      if (!occl) kept.push_back(c[i].id);
    }
    for (size_t k = 0; k < kept.size(); ++k) out[k] = kept[k];
    pcnt[b] = uint32_t(kept.size());
  }
  if (st) {
    st->prune_dist += tot_dist;
    st->prune_pairs += tot_pairs;
  }
}

// add_reverse (Gbuilder.cuh:1020-1038) for every kept edge b -> c, as a CSR
// keyed by target c whose segments list sources in ascending id (C2).  The
// merge step truncates each segment to the cap.  off has n_base+1 entries;
// srcs must hold sum(pcnt) entries.
void orc_reverse_csr(const uint32_t* pruned, const uint32_t* pcnt,
                     uint32_t n_base, uint32_t mp, uint32_t* off,
                     uint32_t* srcs) {
  std::vector<uint32_t> deg(n_base, 0);
  for (uint32_t b = 0; b < n_base; ++b)
    for (uint32_t j = 0; j < pcnt[b]; ++j) {
      uint32_t c = pruned[size_t(b) * mp + j];
      if (c < n_base) ++deg[c];
    }
  off[0] = 0;
  for (uint32_t c = 0; c < n_base; ++c) off[c + 1] = off[c] + deg[c];
  std::vector<uint32_t> cur(off, off + n_base);
  for (uint32_t b = 0; b < n_base; ++b)  // ascending b -> ascending segments
    for (uint32_t j = 0; j < pcnt[b]; ++j) {
      uint32_t c = pruned[size_t(b) * mp + j];
      if (c < n_base) srcs[cur[c]++] = b;
    }
}

// K3 merge_and_sort + fusion_merge_sort_prune_kernel (Gbuilder.cuh:1433-1601):
// candidates = reverse list (first `cap` sources of the CSR segment) followed
// by the pruned list; distances, stable sort, list[0] kept unconditionally,
// strict RNG to `deg`, optional non-strict fill (phase 1), tomb fill.
void orc_merge(const float* base, uint32_t n_base, uint32_t dim,
               const uint32_t* off, const uint32_t* srcs, uint32_t cap,
               const uint32_t* pruned, const uint32_t* pcnt, uint32_t mp,
               uint32_t deg, int strict, uint32_t* out, orc_stats* st) {
  uint64_t tot_dist = 0, tot_pairs = 0, overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tot_dist, tot_pairs, overflow)
  for (int64_t ci = 0; ci < int64_t(n_base); ++ci) {
    uint32_t c = uint32_t(ci);
    uint32_t* o = out + size_t(c) * deg;
    uint32_t seg = off[c + 1] - off[c];
    uint32_t r = std::min(seg, cap);
    if (seg > cap) ++overflow;
    uint32_t p = pcnt[c];
    uint32_t n = r + p;
    for (uint32_t k = 0; k < deg; ++k) o[k] = kTomb;
    if (n == 0) continue;
    const float* xc = base + size_t(c) * dim;
    std::vector<Cand> l(n);
    for (uint32_t j = 0; j < r; ++j) l[j].id = srcs[off[c] + j];
    for (uint32_t j = 0; j < p; ++j) l[r + j].id = pruned[size_t(c) * mp + j];
    for (uint32_t j = 0; j < n; ++j)
      l[j].d = chain_dist(xc, base + size_t(l[j].id) * dim, dim);
    tot_dist += n;
    stable_by_dist(l);
    std::vector<uint32_t> kept;
    kept.reserve(deg);
    kept.push_back(l[0].id);
    for (uint32_t i = 1; i < n && kept.size() < deg; ++i) {
      bool occl;
      tot_pairs += occluded_by(base, dim, l[i], kept, &occl);
      if (!occl) kept.push_back(l[i].id);
    }
    uint32_t k = uint32_t(kept.size());
    if (!strict && k < deg) {
      // Gbuilder.cuh:1559-1594: mark the first list slot matching each kept
      // id, then append unmarked entries in sorted order.
      std::vector<char> added(n, 0);
      for (uint32_t a = 0; a < k; ++a)
        for (uint32_t j = 0; j < n; ++j)
          if (l[j].id == kept[a]) {
            added[j] = 1;
            break;
          }
      for (uint32_t i = 0; i < n && kept.size() < deg; ++i) {
        if (added[i]) continue;
        uint32_t id = l[i].id;
        if (id != kTomb && id != c && id < n_base) kept.push_back(id);
      }
    }
    for (size_t a = 0; a < kept.size(); ++a) o[a] = kept[a];
  }
  if (st) {
    st->merge_dist += tot_dist;
    st->merge_pairs += tot_pairs;
    st->rev_overflow += overflow;
  }
}

// K4 link_process_v0_store_base_data (Gsearcher.cuh:896-1066), one query at a
// time.  The visited hashtable (Ghashtable.cuh:55-135) is an exact set of ids
// inserted since the last reset.  collect_top_p_normal (:75-106) takes the
// first Kp unexpanded entries of list[0, Km).  Output row:
// list[0, topk) & 0x7FFFFFFF.  Entries [0, Km) start as (FLT_MAX, tomb).
void orc_search(const float* queries, uint64_t n_query, const float* base,
                uint32_t n_base, uint32_t dim, const uint32_t* graph,
                uint32_t graph_deg, uint32_t Km, uint32_t Kp, uint32_t Kd,
                uint32_t topk, uint32_t reset_iter, uint32_t entry_seed,
                uint32_t max_iters, uint32_t* out, orc_stats* st) {
  const uint32_t len = Km + Kp * Kd;
  uint64_t tot_dist = 0, tot_iters = 0, capped = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : tot_dist, tot_iters, capped)
  for (int64_t qi = 0; qi < int64_t(n_query); ++qi) {
    const float* q = queries + size_t(qi) * dim;
    std::vector<uint32_t> ids(len);
    std::vector<float> ds(len);
    std::vector<uint32_t> order(len);
    std::vector<uint32_t> tid(len);
    std::vector<float> tds(len);
    std::vector<uint32_t> visited;  // sorted vector as an exact set
    auto sort_list = [&]() {
      for (uint32_t i = 0; i < len; ++i) order[i] = i;
      std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return search_key(ds[a], ids[a]) < search_key(ds[b], ids[b]);
      });
      for (uint32_t i = 0; i < len; ++i) {
        tid[i] = ids[order[i]];
        tds[i] = ds[order[i]];
      }
      ids.swap(tid);
      ds.swap(tds);
    };
    auto reset_visited = [&]() {
      visited.clear();
      for (uint32_t i = 0; i < Km; ++i) visited.push_back(ids[i] & kMaskedTomb);
      std::sort(visited.begin(), visited.end());
      visited.erase(std::unique(visited.begin(), visited.end()), visited.end());
    };
    auto test_and_set = [&](uint32_t id) {
      auto it = std::lower_bound(visited.begin(), visited.end(), id);
      if (it != visited.end() && *it == id) return false;
      visited.insert(it, id);
      return true;
    };
    for (uint32_t i = 0; i < Km; ++i) {
      ids[i] = kTomb;
      ds[i] = FLT_MAX;
    }
    for (uint32_t i = 0; i < Kp * Kd; ++i) {
      uint32_t id = uint32_t((uint64_t(entry_seed) + i) % n_base);
      ids[Km + i] = id;
      ds[Km + i] = search_dist(q, base + size_t(id) * dim, dim);
      ++tot_dist;
    }
    sort_list();
    reset_visited();
    uint32_t iter = 0, total_iter = 0;
    std::vector<uint32_t> explore(Kp);
    while (true) {
      // collect_top_p_normal
      uint32_t found = 0;
      for (uint32_t i = 0; i < Km && found < Kp; ++i) {
        if ((ids[i] & kMsb) == 0) {
          explore[found++] = ids[i];
          ids[i] |= kMsb;
        }
      }
      if (found == 0) break;
      if (total_iter >= max_iters) {
        ++capped;
        break;
      }
      for (uint32_t e = found; e < Kp; ++e) explore[e] = kTomb;
      for (uint32_t e = 0; e < Kp; ++e)
        for (uint32_t j = 0; j < Kd; ++j)
          ids[Km + e * Kd + j] =
              explore[e] == kTomb ? kTomb
                                  : graph[size_t(explore[e]) * graph_deg + j];
      for (uint32_t t = 0; t < Kp * Kd; ++t) {
        uint32_t id = ids[Km + t];
        if (id != kTomb && test_and_set(id)) {
          ds[Km + t] = search_dist(q, base + size_t(id) * dim, dim);
          ++tot_dist;
        } else {
          ds[Km + t] = FLT_MAX;
        }
      }
      sort_list();
      ++total_iter;
      if (++iter == reset_iter) {
        reset_visited();
        iter = 0;
      }
    }
    tot_iters += total_iter;
    uint32_t* o = out + size_t(qi) * topk;
    for (uint32_t i = 0; i < topk; ++i) o[i] = ids[i] & kMaskedTomb;
  }
  if (st) {
    st->search_dist += tot_dist;
    st->search_iters += tot_iters;
    st->search_capped += capped;
  }
}

struct orc_params {
  uint32_t n_base, dim, n_query, gt_k;
  uint32_t proj_deg, proj_pr, proj_rev;
  uint32_t M, link_rev, Km, Kp, Kd, topk, reset_iter, rounds, entry_seed,
      max_iters;
};

// The whole GPU stage (tests/test_t2iend.cu:968-1155): phase-1 projection
// (match, prune 128->15, reverse cap 111, merge to 16 non-strict), then
// `rounds` link rounds (search over the current graph, prune topk->M,
// reverse cap 126-M, merge to M strict).
int orc_gpu_stage(const orc_params* p, const float* base, const uint32_t* gt,
                  uint32_t* out_proj, uint32_t* out_link, orc_stats* st) {
  const uint32_t N = p->n_base;
  const uint32_t mp_max = std::max(p->proj_pr, p->M);
  std::vector<uint32_t> winner(N), rows(size_t(N) * p->gt_k);
  std::vector<uint32_t> pruned(size_t(N) * mp_max), pcnt(N), off(N + 1);
  std::vector<uint32_t> srcs(size_t(N) * mp_max);
  orc_match_winner(gt, p->n_query, p->gt_k, N, winner.data());
  orc_top1_rows(gt, p->gt_k, winner.data(), N, rows.data());
  orc_prune(base, N, p->dim, rows.data(), p->gt_k, p->proj_pr, pruned.data(),
            pcnt.data(), st);
  orc_reverse_csr(pruned.data(), pcnt.data(), N, p->proj_pr, off.data(),
                  srcs.data());
  orc_merge(base, N, p->dim, off.data(), srcs.data(), p->proj_rev,
            pruned.data(), pcnt.data(), p->proj_pr, p->proj_deg, 0, out_proj,
            st);
  std::vector<uint32_t> res(size_t(N) * p->topk);
  const uint32_t* graph = out_proj;
  uint32_t gdeg = p->proj_deg;
  for (uint32_t r = 0; r < p->rounds; ++r) {
    orc_search(base, N, base, N, p->dim, graph, gdeg, p->Km, p->Kp, p->Kd,
               p->topk, p->reset_iter, p->entry_seed, p->max_iters, res.data(),
               st);
    orc_prune(base, N, p->dim, res.data(), p->topk, p->M, pruned.data(),
              pcnt.data(), st);
    orc_reverse_csr(pruned.data(), pcnt.data(), N, p->M, off.data(),
                    srcs.data());
    orc_merge(base, N, p->dim, off.data(), srcs.data(), p->link_rev,
              pruned.data(), pcnt.data(), p->M, p->M, 1, out_link, st);
    graph = out_link;
    gdeg = p->M;
  }
  return 0;
}

// Node-subset variants for timing on a bounded sample: the same per-node
// work as orc_prune / orc_merge, for nodes[0..n_nodes) only.  rows_s holds
// the sample's candidate rows (n_nodes x D); outputs are per sample entry.
void orc_prune_nodes(const float* base, uint32_t n_base, uint32_t dim,
                     const uint32_t* nodes, uint32_t n_nodes,
                     const uint32_t* rows_s, uint32_t D, uint32_t mp,
                     uint32_t* pruned_s, uint32_t* pcnt_s) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t si = 0; si < int64_t(n_nodes); ++si) {
    const uint32_t b = nodes[si];
    const uint32_t* row = rows_s + size_t(si) * D;
    const float* xb = base + size_t(b) * dim;
    std::vector<Cand> c(D);
    for (uint32_t j = 0; j < D; ++j) {
      uint32_t id = row[j];
      c[j].id = id;
      c[j].d = (id == kTomb || id == b || id == kMaskedTomb || id >= n_base)
                   ? FLT_MAX
                   : chain_dist(xb, base + size_t(id) * dim, dim);
    }
    stable_by_dist(c);
    pcnt_s[si] = 0;
    if (is_stop_id(c[0].id, n_base)) continue;
    std::vector<uint32_t> kept{c[0].id};
    for (uint32_t i = 1; i < D && kept.size() < mp; ++i) {
      if (is_stop_id(c[i].id, n_base)) break;
      bool occl;
      occluded_by(base, dim, c[i], kept, &occl);
      if (!occl) kept.push_back(c[i].id);
    }
    for (size_t k = 0; k < kept.size(); ++k) pruned_s[size_t(si) * mp + k] = kept[k];
    pcnt_s[si] = uint32_t(kept.size());
  }
}

void orc_merge_nodes(const float* base, uint32_t n_base, uint32_t dim,
                     const uint32_t* nodes, uint32_t n_nodes,
                     const uint32_t* off, const uint32_t* srcs, uint32_t cap,
                     const uint32_t* pruned, const uint32_t* pcnt, uint32_t mp,
                     uint32_t deg, int strict, uint32_t* out_s) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t si = 0; si < int64_t(n_nodes); ++si) {
    const uint32_t c = nodes[si];
    uint32_t* o = out_s + size_t(si) * deg;
    const uint32_t seg = off[c + 1] - off[c];
    const uint32_t r = std::min(seg, cap), p = pcnt[c], n = r + p;
    for (uint32_t k = 0; k < deg; ++k) o[k] = kTomb;
    if (n == 0) continue;
    const float* xc = base + size_t(c) * dim;
    std::vector<Cand> l(n);
    for (uint32_t j = 0; j < r; ++j) l[j].id = srcs[off[c] + j];
    for (uint32_t j = 0; j < p; ++j) l[r + j].id = pruned[size_t(c) * mp + j];
    for (uint32_t j = 0; j < n; ++j)
      l[j].d = chain_dist(xc, base + size_t(l[j].id) * dim, dim);
    stable_by_dist(l);
    std::vector<uint32_t> kept{l[0].id};
    for (uint32_t i = 1; i < n && kept.size() < deg; ++i) {
      bool occl;
      occluded_by(base, dim, l[i], kept, &occl);
      if (!occl) kept.push_back(l[i].id);
    }
    if (!strict && kept.size() < deg) {
      std::vector<char> added(n, 0);
      const size_t k = kept.size();
      for (size_t a = 0; a < k; ++a)
        for (uint32_t j = 0; j < n; ++j)
          if (l[j].id == kept[a]) {
            added[j] = 1;
            break;
          }
      for (uint32_t i = 0; i < n && kept.size() < deg; ++i)
        if (!added[i] && l[i].id != kTomb && l[i].id != c && l[i].id < n_base)
          kept.push_back(l[i].id);
    }
    for (size_t a = 0; a < kept.size(); ++a) o[a] = kept[a];
  }
}

int orc_max_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
