/* paragraph_host.h — C ABI of the host-side stages that complete a
 * ParaGraph index around the GPU stage (include/paragraph.h).
 *
 * In the reference these run on a CPU thread beside the GPU stage and after
 * it (tests/test_t2iend.cu:888-962 and :1228-1258; `compare` :31-78, RNGPrune
 * :99-177):
 *   MatchNN     -> only its entry point survives (the lists it returns are
 *                  empty: its graph-building loop is commented out,
 *                  tests/test_t2iend.cu:179-321);
 *   MatchSup    -> one supply graph per gt column 1..rounds (:365-488);
 *   FusionNN    -> the bipartite graph from the supply graphs (:490-537);
 *   FusionFinal -> the index lists from the bipartite graph, the GPU link
 *                  rows and the GPU projection rows (:539-613);
 *   SaveGraph   -> the index file (:615-633).
 * The reference exports none of these (they live in a gtest file), so each
 * entry point names the routine it replaces.
 *
 * Distances are the reference's CPU `compare` (tests/test_t2iend.cu:31-78):
 * -dot with 16 FMA lanes over blocks of 16, halved to 8 lanes (+ one block
 * of 8), to 4 lanes (+ one block of 4), a masked remainder, then two
 * horizontal adds.  Where the reference depends on thread scheduling or on
 * hash-set iteration order, these rules fix the result:
 *   C4  entry point: the most frequent id of MatchNN's scan, ties to the
 *       smallest id (reference: std::sort, not stable, :217-221);
 *   C6  a MatchSup reverse list over its cap (128 - M_nn) keeps the sources
 *       of the lowest query ids (the reference's order with one thread; with
 *       more, its dynamic schedule decides, :429-440);
 *   C7  a MatchSup list of at most M_nn merged ids is emitted in ascending
 *       id (reference: unordered_set iteration order, :456-461; only the set
 *       reaches FusionNN, which de-duplicates and sorts);
 *   C8  dim % 4 != 0: the remainder of `compare` uses elements 1..r of the
 *       last partial block of 4, as the reference's mask does (:64-70); the
 *       element past the row end that this mask reaches is taken as 0 (the
 *       reference reads the next row there).
 *
 * Lists are CSR: off[n + 1] offsets into ids[].  Calls return 0 or a
 * negative pg_status (paragraph.h); pg_host_last_error() gives the message.
 * `threads` = OpenMP threads (0 = the runtime's default; the reference caps
 * at 64, tests/test_t2iend.cu:739-741).
 */
#ifndef PARAGRAPH_HOST_H_
#define PARAGRAPH_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pg_host_params {
  uint32_t n_base;       /* N base vectors */
  uint32_t dim;          /* vector dimension */
  uint32_t n_query;      /* gt rows */
  uint32_t gt_k;         /* gt row width (128; ids >= N end a row) */
  uint32_t M_nn;         /* supply / bipartite degree (40; FusionNN keeps up
                            to 2 * M_nn, tests/test_t2iend.cu:527,915,939) */
  uint32_t rounds;       /* MatchSup rounds, gt columns 1..rounds (6,
                            MATCH_ROUNDS, tests/test_t2iend.cu:637-650) */
  uint32_t M_final;      /* FusionFinal degree (70, :1230) */
  uint32_t threads;      /* OpenMP threads, 0 = default */
} pg_host_params;

/* Fills the reference's values: M_nn 40, rounds 6, M_final 70, gt_k 128. */
int pg_host_params_init(pg_host_params* p, uint32_t n_base, uint32_t dim,
                        uint32_t n_query);

/* MatchNN (tests/test_t2iend.cu:179-221): the entry point written into the
 * index header (rule C4). */
int pg_host_entry_point(const pg_host_params* p, const uint32_t* gt,
                        uint32_t* ep);

/* MatchSup round `topn` (tests/test_t2iend.cu:365-488): supply lists of at
 * most M_nn ids per base.  off: n_base + 1; ids: n_base * M_nn. */
int pg_host_match_sup(const pg_host_params* p, const float* base,
                      const uint32_t* gt, uint32_t topn, uint32_t* off,
                      uint32_t* ids);

/* MatchSup rounds 1..rounds and FusionNN (tests/test_t2iend.cu:913-941,
 * 490-537): the bipartite lists, at most 2 * M_nn ids per base.
 * off: n_base + 1; ids: n_base * 2 * M_nn. */
int pg_host_bipartite(const pg_host_params* p, const float* base,
                      const uint32_t* gt, uint32_t* off, uint32_t* ids);

/* FusionFinal (tests/test_t2iend.cu:539-613): per base, its bipartite list
 * followed by the RNG-pruned (M_final, strict) union of bipartite, link and
 * projection candidates that are not already in it, up to M_final entries in
 * all.  As in the reference the room left is computed in unsigned
 * arithmetic, so a bipartite list longer than M_final takes every pruned
 * candidate (rows up to 2 * M_nn + M_final).  link: n_base x M_link rows,
 * proj: n_base x proj_deg rows (the GPU stage's outputs; ids >= N are
 * skipped).  off: n_base + 1; ids: n_base * pg_host_final_width(p). */
uint32_t pg_host_final_width(const pg_host_params* p);
int pg_host_fusion_final(const pg_host_params* p, const float* base,
                         const uint32_t* bi_off, const uint32_t* bi_ids,
                         const uint32_t* link, uint32_t M_link,
                         const uint32_t* proj, uint32_t proj_deg,
                         uint32_t* off, uint32_t* ids);

/* SaveGraph (tests/test_t2iend.cu:615-633): uint32 ep, uint32 n, then per
 * base uint32 count and count ids. */
int pg_host_save_index(const char* path, uint32_t ep, uint32_t n,
                       const uint32_t* off, const uint32_t* ids);

/* Vector and gt files.  Big-ANN .fbin / .ibin: uint32 rows, uint32 cols,
 * then rows x cols 4-byte elements (float32 / int32), row-major (README.md:
 * 15-21); the reference's loader reads the same arrays with no header
 * (include/utils.hpp:18-34, after transform_data_to_gpu.ipynb strips it).
 * pg_host_bin_shape reads the header.  pg_host_read_bin reads rows
 * [row0, row0 + nrows) into out (nrows x cols elements), so a large file
 * can be streamed chunk by chunk into pinned staging memory
 * (pg_host_alloc, paragraph.h); headered = 0 reads a headerless file of
 * `cols` columns, headered = 1 checks cols against the header (0 = take
 * the header's). */
int pg_host_bin_shape(const char* path, uint32_t* rows, uint32_t* cols);
int pg_host_read_bin(const char* path, int headered, uint32_t cols,
                     uint64_t row0, uint64_t nrows, void* out);

/* gt rows re-padded (or truncated) to out_cols columns with -1, as
 * transform_gt_to_128.ipynb pads the reference's gt to 128. */
int pg_host_pad_gt(const int32_t* in, uint64_t rows, uint32_t in_cols,
                   uint32_t out_cols, int32_t* out);

const char* pg_host_last_error(void);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* PARAGRAPH_HOST_H_ */
