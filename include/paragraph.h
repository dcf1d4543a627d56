/* paragraph.h — C ABI of the MI355X-native ParaGraph GPU index-build stage.
 *
 * Replaces the reference's GPU stage: the 12 kernel launches on
 * compute_stream plus the H2D/D2H copies in tests/test_t2iend.cu:850-1155
 * (kernels in include/Gbuilder.cuh and include/Gsearcher.cuh).  The
 * reference exposes no library entry point for this stage (its
 * libGGidxbuildCUDA exports only test(), src/Gbuilder.cu:1-8; its Builder
 * class is a commented-out stub, include/Builder.hpp:6-41), so each entry
 * point below names the reference code it stands in for.
 *
 * Conventions
 *  - Plain C types only; ids are uint32_t, vectors float32 row-major.
 *  - tomb = 0xFFFFFFFF (int32 -1) marks empty slots, as in the reference.
 *  - Every call returns 0 on success or a negative pg_status; nothing
 *    calls exit().  pg_last_error() gives the message.
 *  - A context owns its HIP stream, its workspace and its output staging;
 *    one context per host thread.  Work is enqueued on that stream only:
 *    device buffers passed to a call (inputs, and outputs the caller
 *    pre-filled) must be complete when the call is made (synchronize the
 *    stream that produced them), and outputs are valid after pg_wait /
 *    pg_sync.  Calls do not spawn CPU threads, so they
 *    may run beside an OpenMP CPU stage (tests/test_t2iend.cu:888-962).
 */
#ifndef PARAGRAPH_H_
#define PARAGRAPH_H_

#include <stdint.h>

#include "paragraph_host.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PG_TOMB 0xFFFFFFFFu

typedef enum pg_status {
  PG_OK = 0,
  PG_ERR_INVALID = -1,  /* parameter or size check failed */
  PG_ERR_OOM = -2,      /* device allocation failed */
  PG_ERR_HIP = -3,      /* a HIP runtime call or kernel failed */
  PG_ERR_BUSY = -4,     /* pg_query: work still in flight */
  PG_ERR_NO_DEVICE = -5 /* no usable GPU */
} pg_status;

typedef enum pg_precision {
  PG_FP32 = 0,     /* plain FP32 path: reproduces the reference ids exactly */
  PG_BF16 = 1,     /* BF16 screening on MFMA, FP32 re-rank before pruning */
  PG_FP8E4M3 = 2   /* FP8 e4m3 screening with per-vector scales */
} pg_precision;

/* Build parameters.  The reference fixes these as constexpr template
 * arguments (tests/test_t2iend.cu:670-727); here they are runtime values.
 * pg_params_init() fills the mapping used by this project for a given
 * (n_base, dim, M, L). */
typedef struct pg_params {
  uint32_t n_base;      /* N base vectors (< 2^31: the MSB is the expanded flag,
                           Gsearcher.cuh:906) */
  uint32_t dim;         /* vector dimension */
  uint32_t n_query;     /* gt rows (training queries) */
  uint32_t gt_k;        /* gt row width (128, -1 padded) */
  uint32_t proj_deg;    /* phase-1 projection degree (16) */
  uint32_t proj_pr;     /* phase-1 pruned cap (15) */
  uint32_t proj_rev;    /* phase-1 reverse cap (111) */
  uint32_t M;           /* link degree = round pruned cap */
  uint32_t link_rev;    /* round reverse cap (126 - M) */
  uint32_t Km;          /* search top-list length (beam L) */
  uint32_t Kp;          /* nodes expanded per iteration */
  uint32_t Kd;          /* neighbours read per expanded node */
  uint32_t topk;        /* search output width (2L) */
  uint32_t reset_iter;  /* visited-table reset period (15) */
  uint32_t table_size;  /* visited-table slots in LDS; 0 = auto (the most ids
                           one reset period can insert, Km + reset_iter*Kp*Kd,
                           plus 1/16 slack).  The reference uses 4096
                           (tests/test_t2iend.cu:694); the size changes speed
                           only, never results. */
  uint32_t rounds;      /* link rounds (2) */
  uint32_t entry_seed;  /* entry points (seed + i) mod N (3030517) */
  uint32_t max_iters;   /* search iteration watchdog (not in the reference) */
  uint32_t precision;   /* pg_precision */
  uint32_t rerank;      /* screened paths: the search keeps its first rerank
                           entries (screened order) and the prune re-ranks
                           them in FP32 (0 = topk) */
  uint32_t tune_flags;  /* perf-only knobs, 0 = defaults.  bits 0-4: search
                           gather batch (FP32: candidate pairs in flight;
                           BF16: 16-candidate groups); bit 5 (0x20):
                           persistent search grid fed by a device work
                           counter; bit 6 (0x40): capture the pipeline in a
                           hipGraph once per buffer set and replay it (no
                           per-stage times in that mode); bit 7 (0x80):
                           split-BF16 screen in the screened merge and
                           candidate-major prune (tighter bound, fewer
                           exact fallbacks; the kept-major sweep keeps the
                           plain screen); bit 8
                           (0x100): prune the link rows candidate by
                           candidate (the reference's loop order) instead of
                           kept-major sweeps (same rows); bit 9 (0x200):
                           the persistent search takes its queries in a
                           locality order (nodes grouped by nearest pivot
                           row, one queue segment per XCD), and the
                           link-row prune its nodes; bit 10 (0x400),
                           with bit 9: the link-row merge takes its nodes
                           in that order too (same rows); bit 11 (0x800):
                           two waves per search query, the second sharing
                           its distance evaluations (same rows) */
  uint32_t reserved[7];
} pg_params;

/* Counters of one build (same meaning as the oracle's). */
typedef struct pg_stats {
  uint64_t prune_dist;    /* node->candidate distances in prune steps */
  uint64_t prune_pairs;   /* occlusion pair distances in prune steps (the
                             screened link-row sweep counts its screened
                             pair decisions instead) */
  uint64_t merge_dist;    /* node->candidate distances in merge steps */
  uint64_t merge_pairs;   /* occlusion pair distances in merge steps */
  uint64_t search_dist;   /* visited-miss distance evaluations */
  uint64_t search_iters;  /* search iterations */
  uint64_t search_capped; /* queries stopped by the watchdog */
  uint64_t rev_overflow;  /* targets whose reverse list exceeded the cap */
  uint64_t screen_fallback; /* screened paths: exact re-evaluations */
  uint64_t reserved[7];
  float stage_ms[16];     /* per-stage device time of the last build (see
                             pg_stage_name) */
  uint32_t n_stages;
} pg_stats;

typedef struct pg_ctx pg_ctx;

/* Fills *p with the project's mapping of (N, dim, M, L) onto the reference
 * knobs: link_rev = 126-M, Km = L, Kd = 16, Kp = ceil(L/16), topk = 2L,
 * reset_iter 15, table auto, 2 rounds, seed 3030517, projection 16/15/111
 * (reference values: tests/test_t2iend.cu:676-707); tune_flags = persistent
 * search grid. */
int pg_params_init(pg_params* p, uint32_t n_base, uint32_t dim,
                   uint32_t n_query, uint32_t M, uint32_t L);

/* Checks a parameter set without touching the GPU (used by pg_create). */
int pg_params_check(const pg_params* p, char* msg, uint32_t msg_len);

/* Creates a context on `device` and allocates its workspace for *p
 * (replaces the cudaMalloc block, tests/test_t2iend.cu:781-797).  Returns
 * NULL on failure; pg_last_error(NULL) then holds the reason. */
pg_ctx* pg_create(int device, const pg_params* p);
void pg_destroy(pg_ctx* ctx);

/* Whole GPU stage from host buffers: H2D of base (N x dim) and gt
 * (n_query x gt_k), the 12-launch pipeline, D2H of the projection
 * (N x proj_deg) and link (N x M) rows (tests/test_t2iend.cu:858-1155).
 * The inputs stream in on the context's copy stream in about 8 chunks each
 * (256 KiB to 64 MiB per chunk), gt first: the match runs on each gt chunk and the screened
 * copies of each base chunk are made as it lands; the projection rows
 * stream out while the link rounds run.  Host buffers from pg_host_alloc
 * (pinned) copy at full link speed; pageable buffers work too, staged by
 * the runtime.  Returns after enqueueing; outputs are valid once pg_wait()
 * returns 0, and the input buffers must stay unchanged until then. */
int pg_gpu_stage_async(pg_ctx* ctx, const float* base, const uint32_t* gt,
                       uint32_t* out_proj, uint32_t* out_link);

/* Same pipeline on device-resident buffers (no copies). */
int pg_gpu_stage_device_async(pg_ctx* ctx, const float* d_base,
                              const uint32_t* d_gt, uint32_t* d_proj,
                              uint32_t* d_link);

/* Pinned (page-locked) host memory for the inputs and outputs of
 * pg_gpu_stage_async: copies from and to it run at full link speed and
 * overlap the kernels.  Replaces the pageable std::vector buffers the
 * reference copies from (tests/test_t2iend.cu:735-760, 858-873).  Free
 * with pg_host_free. */
int pg_host_alloc(uint64_t bytes, void** out);
int pg_host_free(void* ptr);

int pg_wait(pg_ctx* ctx);   /* blocks until the last enqueued work ends */
int pg_query(pg_ctx* ctx);  /* PG_OK if done, PG_ERR_BUSY if in flight */
int pg_get_stats(pg_ctx* ctx, pg_stats* out);  /* waits, then copies */
int pg_reset_stats(pg_ctx* ctx);
/* Counters accumulated during pipeline stage `stage` (0 .. n_stages-1) of
 * the last build, in pg_stats field order (prune_dist, prune_pairs,
 * merge_dist, merge_pairs, search_dist, search_iters, search_capped,
 * rev_overflow, screen_fallback, 7 reserved). */
int pg_get_stage_counters(pg_ctx* ctx, uint32_t stage, uint64_t* out16);
const char* pg_last_error(const pg_ctx* ctx);
const char* pg_stage_name(uint32_t i);

/* ---- Single stages on device buffers (for parity tests and callers that
 * stitch their own pipeline).  All run on the context's stream and use its
 * n_base / dim / search parameters. ---- */

/* match_top1_kernel_v0 (Gbuilder.cuh:104-135): winner[b] = lowest query
 * whose gt row starts with b, else tomb. */
int pg_run_match(pg_ctx* ctx, const uint32_t* d_gt, uint32_t n_query,
                 uint32_t gt_k, uint32_t* d_winner);

/* fusion_prune_reverse_kernel_v0 (Gbuilder.cuh:1105-1244), pruned part.
 * Candidates per node come from d_rows (N x D) or, when d_rows is NULL,
 * from the top-1 row of the winner (gt[w][1..gt_k-2], two tombs; D = gt_k).
 * Writes pruned (N x mp) and counts (N). */
int pg_run_prune(pg_ctx* ctx, const float* d_base, const uint32_t* d_rows,
                 uint32_t D, const uint32_t* d_gt, const uint32_t* d_winner,
                 uint32_t gt_k, uint32_t mp, uint32_t* d_pruned,
                 uint32_t* d_pcnt);

/* add_reverse (Gbuilder.cuh:1020-1038) as a CSR by target: d_off (N+1),
 * d_srcs (sum of counts).  Segment order is unspecified; the merge step
 * selects the cap smallest sources in ascending order. */
int pg_run_reverse(pg_ctx* ctx, const uint32_t* d_pruned,
                   const uint32_t* d_pcnt, uint32_t mp, uint32_t* d_off,
                   uint32_t* d_srcs);

/* merge_and_sort + fusion_merge_sort_prune_kernel (Gbuilder.cuh:1433-1601).
 * Writes N x deg rows. */
int pg_run_merge(pg_ctx* ctx, const float* d_base, const uint32_t* d_off,
                 const uint32_t* d_srcs, uint32_t cap,
                 const uint32_t* d_pruned, const uint32_t* d_pcnt,
                 uint32_t mp, uint32_t deg, int strict, uint32_t* d_out);

/* link_process_v0_store_base_data (Gsearcher.cuh:896-1066), generalised
 * to queries that are not base rows (the unused search_shared_hashtable,
 * Gsearcher.cuh:108-296, is the reference model for that).  Writes
 * n_query x topk ids (& 0x7FFFFFFF). */
int pg_run_search(pg_ctx* ctx, const float* d_queries, uint32_t n_query,
                  const float* d_base, const uint32_t* d_graph,
                  uint32_t graph_deg, uint32_t* d_out);

/* Distances of (query, base id) pairs, n_query x k ids (k <= 1024), written
 * as n_query x k floats (NaN for ids >= n_base).  kind 0: the search
 * stage's evaluator at the context's precision (FP32: the per-lane chains
 * and 16/8/4/2/1 tree of Gsearcher.cuh:1022-1031; BF16 / FP8: the MFMA
 * screening distance); kind 1: the FP32 chain of the prune and merge stages
 * (Gbuilder.cuh:785-797).  Measurement hook for the screening error; the
 * reference has no separate entry for it. */
int pg_run_pair_distances(pg_ctx* ctx, const float* d_queries,
                          uint32_t n_query, const float* d_base,
                          const uint32_t* d_ids, uint32_t k, int kind,
                          float* d_out);

/* The host stages of the full build on the GPU: MatchSup rounds
 * 1..hp->rounds, FusionNN and FusionFinal (tests/test_t2iend.cu:365-613),
 * giving the lists of pg_host_bipartite + pg_host_fusion_final
 * (paragraph_host.h) id for id, under the same rules C6-C8.  Device
 * buffers: base (N x dim), gt (N x gt_k; needs n_query == n_base), proj
 * (N x proj_deg) and link (N x M_link) -- the GPU stage's outputs; out:
 * off (N + 1) and ids (N x pg_host_final_width(hp)), CSR.  hp->n_base and
 * dim must be the context's.  Limits: gt_k <= 256, M_nn <= 64,
 * rounds * M_nn <= 256, 2 M_nn + M_link + proj_deg <= 256, M_final <= 128.
 * The entry point stays on the host (pg_host_entry_point: the reference's
 * scan is sequential).  The workspace (about (256 + 5 M_nn + rounds *
 * (M_nn + 1) + M_final) x 4 bytes per base) is allocated on the first call
 * and kept by the context. */
int pg_run_fusion(pg_ctx* ctx, const pg_host_params* hp, const float* d_base,
                  const uint32_t* d_gt, const uint32_t* d_proj,
                  uint32_t proj_deg, const uint32_t* d_link, uint32_t M_link,
                  uint32_t* d_off, uint32_t* d_ids);

/* The two halves of pg_run_fusion, so that the bipartite stage overlaps
 * the GPU stage as the reference's CPU thread does (tests/test_t2iend.cu:
 * 888-962): pg_run_bipartite_async enqueues MatchSup rounds + FusionNN
 * (they read only base and gt) on the context's second stream, ordered
 * after the work already on its main stream; pg_run_fusion_final enqueues
 * FusionFinal on the main stream after both the bipartite stage and what
 * the main stream holds (e.g. pg_gpu_stage_device_async).  hp must have
 * the same shape in both calls.  pg_wait / pg_query cover both streams. */
int pg_run_bipartite_async(pg_ctx* ctx, const pg_host_params* hp,
                           const float* d_base, const uint32_t* d_gt);
int pg_run_fusion_final(pg_ctx* ctx, const pg_host_params* hp,
                        const float* d_base, const uint32_t* d_proj,
                        uint32_t proj_deg, const uint32_t* d_link,
                        uint32_t M_link, uint32_t* d_off, uint32_t* d_ids);

int pg_sync(pg_ctx* ctx);  /* hipStreamSynchronize + error check */

/* The gt the build reads, computed on the GPU: for each query the k <= 256
 * base ids with the largest inner product, scored by the FP32 sequential
 * chain (Gbuilder.cuh:785-797's order), ties to the lower id, -1 past
 * n_base; d_out is n_query x k int32, the layout the reference reads from
 * its gt file after transform_gt_to_128.ipynb (tests/test_t2iend.cu:
 * 735-760).  A BF16 MFMA screen with a rigorous per-pair bound keeps a
 * superset of every top-k, which is re-scored exactly; a query whose
 * bound admits more than 1024 base rows around its k-th score (crowded
 * scores) is computed by an exact pass over all base rows instead, and
 * *n_exact (may be NULL) counts those rows.  dim <= 512.  Runs on `stream`
 * (a hipStream_t, NULL = the null stream) and returns when done; it
 * allocates its workspace (BF16 copies and about 8 KiB of candidates per
 * query, in chunks of 2^20 queries). */
int pg_exact_topk_ip(const float* d_queries, uint32_t n_query,
                     const float* d_base, uint32_t n_base, uint32_t dim,
                     uint32_t k, int32_t* d_out, uint32_t* n_exact,
                     void* stream);

/* Allocation-free variant of pg_run_search.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, serving loops, harnesses that forbid hipMalloc),
 * so it takes no context.  It runs the exact FP32 search of *p
 * (p->precision must be PG_FP32) on the current device, enqueued on
 * `stream` (a hipStream_t; NULL = the null stream), and calls no
 * allocating routine: its work counters and locality order live in a
 * fixed device array of the library (8 MiB + 2.3 KiB per GPU, held until
 * the process ends).  Limit: with the locality order on (tune_flags bits 5
 * and 9) n_query <= 2^20, else PG_ERR_INVALID.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU race: serialise them
 * (one stream, or events).  Errors are reported by pg_last_error(NULL).
 * Outputs equal pg_run_search's for the same params and inputs. */
int pg_search_noalloc(const pg_params* p, const float* d_queries,
                      uint32_t n_query, const float* d_base,
                      const uint32_t* d_graph, uint32_t graph_deg,
                      uint32_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PARAGRAPH_H_ */
