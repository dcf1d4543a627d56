// Host-side launchers for the index-build kernels (pg_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pg {

// Screened-path inputs of the prune / merge kernels (see pg_kernels.hip).
struct ScreenArgs {
  const __bf16* base16;    // bf16(x), rows padded to dpad, zero-filled
  const __bf16* base16lo;  // bf16(x - hi) for the split screen, or nullptr
  const float* norms;      // FP32 row norms |x|
  const float* rnorms;     // |x - bf16(x)| per row (plain screen), or nullptr
  uint32_t dpad;
  float eps_c;             // split screen: |screened - chain| <= eps_c |x||y|
  float gam_a, gam_c;      // plain screen: accumulation / chain factors
};

// Persistent search queue: one counter per segment of the locality order
// (one segment per XCD).
constexpr uint32_t kQueueSegs = 8;
// Locality order: groups per pivot row and the workspace it needs
// (group id per node, then 2 x kOrderGroups counters).
constexpr uint32_t kOrderGroups = 256;
constexpr uint32_t kOrderDims = 64;

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n,
                           unsigned long long v, hipStream_t s);
hipError_t launch_snapshot_stats(const unsigned long long* stats,
                                 unsigned long long* row, hipStream_t s);
hipError_t launch_match(const uint32_t* gt, uint32_t n_query, uint32_t gt_k,
                        uint32_t n_base, uint32_t* winner, hipStream_t s);
// The match over gt rows of queries q0 .. q0 + n_query - 1 (gt points at
// the first of them) into a winner array already filled with tomb.
hipError_t launch_match_rows(const uint32_t* gt, uint32_t q0, uint32_t n_query,
                             uint32_t gt_k, uint32_t n_base, uint32_t* winner,
                             hipStream_t s);
hipError_t launch_prune(const float* base, uint32_t n_base, uint32_t dim,
                        const uint32_t* rows, uint32_t D, const uint32_t* gt,
                        const uint32_t* winner, uint32_t gt_k, uint32_t mp,
                        uint32_t* pruned, uint32_t* pcnt,
                        unsigned long long* stats, const ScreenArgs* sc,
                        bool sweep, const uint32_t* order, hipStream_t s);
hipError_t launch_reverse(const uint32_t* pruned, const uint32_t* pcnt,
                          uint32_t n_base, uint32_t mp, uint32_t* deg,
                          uint32_t* off, uint32_t* cursor, uint32_t* scan_tmp,
                          uint32_t* srcs, hipStream_t s);
hipError_t launch_merge(const float* base, uint32_t n_base, uint32_t dim,
                        const uint32_t* off, const uint32_t* srcs,
                        uint32_t cap, const uint32_t* pruned,
                        const uint32_t* pcnt, uint32_t mp, uint32_t deg,
                        int strict, uint32_t* out, unsigned long long* stats,
                        const ScreenArgs* sc, const uint32_t* order,
                        hipStream_t s);

size_t search_smem_bytes(uint32_t table_size, uint32_t lp2, uint32_t kpkd,
                         uint32_t Kp);
hipError_t launch_prepare_screen(const float* base, uint32_t n, uint32_t dim,
                                 uint32_t dpad, void* out16, void* lo16,
                                 float* norms, float* rnorms, hipStream_t s);
// base16 == nullptr: exact FP32 search; else BF16-screened search over the
// padded BF16 copy (row stride dpad elements).  queue != nullptr: persistent
// grid fed by the device counter *queue (reset by the launcher).
hipError_t launch_prepare_fp8(const float* base, uint32_t n, uint32_t dim,
                              uint32_t dpad, void* out8, float* scale,
                              hipStream_t s);
// Locality order of the nodes: node i joins the group of its nearest pivot
// row (pivots = rows (g * n) / kOrderGroups, squared L2 over the first
// min(dim, kOrderDims) elements, ties to the lower pivot), and `order` lists
// the nodes group by group.  Only the processing order of the search
// queries depends on it, never a result.  group: n words; cnt: 2 x
// kOrderGroups words.
hipError_t launch_locality_order(const float* base, uint32_t n, uint32_t dim,
                                 uint32_t* group, uint32_t* cnt,
                                 uint32_t* order, hipStream_t s);
// base8 != nullptr: FP8-screened search over per-row-scaled e4m3 rows.
// order != nullptr (persistent grid only): queries are taken in that order,
// kQueueSegs segments with a counter each in queue[0 .. kQueueSegs).
// waves: waves per query (1, or 2: a second wave shares the distance
// evaluations of each query; same results).
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
                         hipStream_t s);

// Host stages of the full build on the GPU (MatchSup rounds 1..rounds,
// FusionNN, FusionFinal; the lists of libparagraph_host).  Needs
// n_query == n_base, gt_k <= 256, rounds * M_nn <= 256,
// 2 * M_nn + M_link + proj_deg <= 256, M_final <= 128 (checked by the
// caller).  off_out: n_base + 1; ids_out: n_base * (2 M_nn + M_final).
struct FusionShape {
  uint32_t n_base, dim, gt_k, M_nn, rounds, M_final, M_link, proj_deg;
};
uint64_t fusion_ws_words(const FusionShape& f);
hipError_t launch_fusion(const FusionShape& f, const float* base,
                         const uint32_t* gt, const uint32_t* proj,
                         const uint32_t* link, uint32_t* ws, uint32_t* off_out,
                         uint32_t* ids_out, hipStream_t s);
// The two halves of launch_fusion: MatchSup + FusionNN (base and gt only)
// into the workspace, then FusionFinal from it.
hipError_t launch_fusion_bipartite(const FusionShape& f, const float* base,
                                   const uint32_t* gt, uint32_t* ws,
                                   hipStream_t s);
hipError_t launch_fusion_final(const FusionShape& f, const float* base,
                               const uint32_t* proj, const uint32_t* link,
                               uint32_t* ws, uint32_t* off_out,
                               uint32_t* ids_out, hipStream_t s);

// Exact top-k by inner product (pg_topk.hip): BF16 MFMA screen with the
// per-pair bound of the screened prune over the BF16 copies (q16 / b16,
// rows of dpad elements; FP32 norms and residual norms |x - bf16(x)| per
// row; gam_a / gam_c the accumulation and chain factors), then the
// survivors re-scored with the FP32 chain; out: nq x k ids (score desc, id
// asc), -1 past the survivors.  A query whose bound admits more pairs than
// its buffer holds gets an exact pass over all base rows instead (list: nq
// words, n_list: one word, the number of such rows after the call).
// buf: nq x kTopkBufferPairs pairs, count: nq words.
hipError_t launch_exact_topk(const float* queries, uint32_t nq,
                             const float* base, uint32_t nb, uint32_t dim,
                             uint32_t k, const void* q16, const float* qnorm,
                             const float* qres, const void* b16,
                             const float* bnorm, const float* bres,
                             uint32_t dpad, float gam_a, float gam_c,
                             uint2* buf, uint32_t* count, uint32_t* list,
                             uint32_t* n_list, int32_t* out, hipStream_t s);
constexpr uint32_t kTopkBufferPairs = 1024;

// Pair distances: chain != 0 gives the FP32 prune chain; else the search
// evaluator selected like launch_search (k <= 1024).
hipError_t launch_pair_dist(const float* queries, uint32_t n_query,
                            const float* base, const void* base16,
                            const void* base8, const float* scale8,
                            uint32_t dpad, uint32_t n_base, uint32_t dim,
                            const uint32_t* ids, uint32_t k, int chain,
                            float* out, hipStream_t s);

}  // namespace pg
