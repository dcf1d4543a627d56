// C ABI of the index-build stage (declared in include/paragraph.h).
//
// Host orchestration that replaces the reference's GPU-stage driver
// (tests/test_t2iend.cu:781-1155): one context owns the stream, the
// workspace and the timing events, and enqueues the pipeline
//   match -> prune(top-1 rows) -> reverse -> merge(16, non-strict)
//   then per round: search -> prune(topk) -> reverse -> merge(M, strict)
// Buffers the reference reuses by reinterpret_cast (d_space_128_xx/yy) are
// separate, typed allocations here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "../../include/paragraph.h"
#include "pg_kernels.h"

namespace {

constexpr int kMaxStages = 16;
constexpr uint64_t kStatCount = 256ull * 16ull;
thread_local char g_err[512] = "";

const char* const kStageNames[kMaxStages] = {
    "match",         "prune_proj",     "reverse_proj",  "merge_proj",
    "search_round1", "prune_round1",   "reverse_round1", "merge_round1",
    "search_round2", "prune_round2",   "reverse_round2", "merge_round2",
    "search_round3", "prune_round3",   "reverse_round3", "merge_round3"};

}  // namespace

struct pg_ctx {
  int device = 0;
  pg_params p{};
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t ev[kMaxStages + 1] = {};
  int n_stage_events = 0;
  uint32_t mp_max = 0;
  // workspace
  uint32_t* winner = nullptr;
  uint32_t* pruned = nullptr;
  uint32_t* pcnt = nullptr;
  uint32_t* deg = nullptr;
  uint32_t* cursor = nullptr;
  uint32_t* off = nullptr;
  uint32_t* scan_tmp = nullptr;
  uint32_t* srcs = nullptr;
  uint32_t* res = nullptr;
  void* base16 = nullptr;   // screened paths: BF16 copy of the base
  void* base16lo = nullptr; // split screen: residual BF16 copy
  void* base8 = nullptr;    // FP8 path: per-row scaled e4m3 rows
  float* scale8 = nullptr;  // FP8 path: row scales
  float* norms = nullptr;   // screened paths: FP32 row norms
  float* rnorms = nullptr;  // plain screen: |x - bf16(x)| per row
  uint32_t dpad = 0;        // BF16 row stride (dim rounded up to 32)
  float eps_c = 0.0f;       // split screen bound factor
  float gam_a = 0.0f, gam_c = 0.0f;  // plain screen: MFMA / chain error
  unsigned long long* stats = nullptr;
  unsigned long long* stage_cnt = nullptr;  // (kMaxStages+1) x 16 snapshots
  uint32_t* queue = nullptr;                // persistent-search work counters
  uint32_t* order = nullptr;   // locality order of the search queries
  uint32_t* ogroup = nullptr;  // its pivot group per node
  uint32_t* ocnt = nullptr;    // its group counters
  const void* order_rows = nullptr;  // rows and count the order was built
  uint32_t order_n = 0;              // for (reused by the pg_run_* calls)
  uint32_t* fuse_ws = nullptr;  // host stages on the GPU (pg_run_fusion)
  uint64_t fuse_words = 0;
  hipStream_t aux = nullptr;    // second stream: the bipartite stage beside
  hipEvent_t ev_in = nullptr;   // the GPU stage (pg_run_bipartite_async)
  hipEvent_t ev_bip = nullptr;
  bool bip_valid = false;       // the workspace holds bipartite rows for
  pg_host_params bip_hp = {};   // this shape
  bool capturing = false;                   // recording a hipGraph
  hipGraph_t graph = nullptr;               // captured pipeline
  hipGraphExec_t graph_exec = nullptr;
  const void* graph_key[4] = {};            // buffers the graph was built for
  // staging for the host-pointer entry point (allocated on first use)
  float* h_base_dev = nullptr;
  uint32_t* h_gt_dev = nullptr;
  uint32_t* h_proj_dev = nullptr;
  uint32_t* h_link_dev = nullptr;
  // its copy stream and events: inputs stream in chunk by chunk beside the
  // first kernels, the projection rows stream out beside the link rounds
  hipStream_t copy = nullptr;
  hipEvent_t ev_chunk[8] = {};
  hipEvent_t ev_proj = nullptr;
  hipEvent_t ev_out = nullptr;
  char err[512] = "";
};

namespace {

int set_err(pg_ctx* c, int code, const char* fmt, ...) {
  char* dst = c ? c->err : g_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  if (c) std::memcpy(g_err, c->err, sizeof(g_err));
  return code;
}

#define PG_HIP(ctx, call)                                                 \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      return set_err((ctx), PG_ERR_HIP, "%s failed: %s", #call,           \
                     hipGetErrorString(e_));                              \
  } while (0)

template <typename T>
int dev_alloc(pg_ctx* c, T** ptr, uint64_t count, const char* what) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return set_err(c, PG_ERR_OOM, "hipMalloc(%s, %llu bytes): %s", what,
                   (unsigned long long)(count * sizeof(T)),
                   hipGetErrorString(e));
  }
  return PG_OK;
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Visited-table slots actually used: table_size, or (0 = auto) the most ids
// one reset period can insert plus 1/16 slack, rounded up to 64 slots.
// The screened search uses the table as a visited cache (see the kernel); its
// auto size is 4*(Km + Kp*Kd) rounded up to 64 slots.
uint32_t effective_table_size(const pg_params& p) {
  if (p.table_size) return p.table_size;
  if (p.precision != PG_FP32) {
    const uint64_t t = (4ull * (p.Km + uint64_t(p.Kp) * p.Kd) + 63) / 64 * 64;
    return uint32_t(t);
  }
  const uint64_t bound = uint64_t(p.Km) + uint64_t(p.reset_iter) * p.Kp * p.Kd;
  uint64_t t = bound + (bound / 16 > 64 ? bound / 16 : 64);
  t = (t + 63) / 64 * 64;
  return t > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(t);
}

// Perf-only knobs packed in tune_flags (0 = library default):
//   bits 0-4  search gather batch (FP32: candidate pairs in flight, 4/8/16;
//             BF16: 16-candidate groups in flight, 1/2)
//   bit 5     persistent search grid fed by a device work counter
//   bit 6     capture the pipeline in a hipGraph once, replay it per build
//   bit 7     split-BF16 screen in the merge and the candidate-major prune
//             (x = hi + lo, tighter bound, twice the screen bytes); the
//             kept-major sweep screens with the plain bound in every
//             screened context
//   bit 8     link-row prune candidate by candidate (the reference's loop
//             order) instead of the kept-major sweep
//   bit 9     persistent search takes its queries in a locality order
//             (nodes grouped by nearest pivot row, one queue segment per
//             XCD), and the link-row prune its nodes; same rows
//   bit 10    with bit 9, the merge takes its nodes in that order too
//   bit 11    two waves per search query (the second shares the distance
//             evaluations); same rows
constexpr uint32_t kTunePersistent = 0x20u;
constexpr uint32_t kTuneGraph = 0x40u;
constexpr uint32_t kTuneSplit = 0x80u;
constexpr uint32_t kTuneScanPrune = 0x100u;
constexpr uint32_t kTuneOrder = 0x200u;
constexpr uint32_t kTuneOrderMerge = 0x400u;
constexpr uint32_t kTuneWaves2 = 0x800u;
uint32_t tune_gather(const pg_params& p) { return p.tune_flags & 0x1Fu; }
uint32_t tune_waves(const pg_params& p) {
  return (p.tune_flags & kTuneWaves2) ? 2u : 1u;
}

// Bound factor of the BF16 screen against the exact FP32 chain:
// |screened - chain| <= eps * sum|x_i y_i| <= eps * |x| |y| with
//   BF16 rounding of both operands (2u + u^2, u = 2^-8; the products of
//   BF16 values are exact in FP32),
//   FP32 accumulation inside the MFMA over dpad terms (gamma taken with
//   2^-23 to allow any rounding mode; the terms sum to at most
//   (1 + u)^2 sum|x y|),
//   the chain's own FP32 error (gamma over dim terms, u = 2^-24),
// inflated by 1% for the FP32 norms and the comparisons themselves.
float screen_eps_split(uint32_t dim, uint32_t dpad);

float screen_eps(uint32_t dim, uint32_t dpad) {
  const double u = 1.0 / 256.0;
  const double ua = std::ldexp(1.0, -23);
  const double ga = dpad * ua / (1.0 - dpad * ua);
  const double gc = dim * std::ldexp(1.0, -24) / (1.0 - dim * std::ldexp(1.0, -24));
  const double e = (2 * u + u * u) + ga * (1 + u) * (1 + u) + gc;
  return float(e * 1.01);
}

// Split screen: x = hi + lo + r with |hi| <= (1+u)|x|, |lo| <= u(1+u)|x|,
// |r| <= u^2|x| (u = 2^-8).  The three products hi.hi, lo.hi, hi.lo are
// exact in FP32; the dropped terms hi.r', lo.lo', lo.r', r.y are below
// (3u^2 + 4u^3 + 2u^4)|x y|.  The MFMA accumulates 3*dpad terms whose
// magnitudes sum to at most (1+u)^2 (1+2u) sum|x y| (gamma with 2^-23); the
// chain's own error is gamma(dim) with 2^-24.  Inflated by 1%.
float screen_eps_split(uint32_t dim, uint32_t dpad) {
  const double u = 1.0 / 256.0;
  const double na = 3.0 * dpad, ua = std::ldexp(1.0, -23);
  const double ga = na * ua / (1.0 - na * ua);
  const double gc = dim * std::ldexp(1.0, -24) / (1.0 - dim * std::ldexp(1.0, -24));
  const double e = 3 * u * u + 4 * u * u * u + 2 * u * u * u * u +
                   ga * (1 + u) * (1 + u) * (1 + 2 * u) + gc;
  return float(e * 1.01);
}

uint32_t next_pow2(uint32_t x) {
  uint32_t v = 1;
  while (v < x) v <<= 1;
  return v;
}

int record(pg_ctx* c, int idx) {
  if (idx <= kMaxStages) {
    // stage events are not recorded into a graph; counters still are
    if (!c->capturing) PG_HIP(c, hipEventRecord(c->ev[idx], c->stream));
    PG_HIP(c, pg::launch_snapshot_stats(c->stats, c->stage_cnt + idx * 16,
                                        c->stream));
  }
  return PG_OK;
}

#define PG_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != PG_OK) return rc_; \
  } while (0)

pg::ScreenArgs screen_args(const pg_ctx* c) {
  return pg::ScreenArgs{reinterpret_cast<const __bf16*>(c->base16),
                        reinterpret_cast<const __bf16*>(c->base16lo), c->norms,
                        c->base16lo ? nullptr : c->rnorms, c->dpad, c->eps_c,
                        c->gam_a, c->gam_c};
}

// The link-row sweep always screens with the plain BF16 copy and the
// residual-norm bound (its exact chains run as quad chains): a split-screen
// context keeps the split for the merge and the candidate-major prune.
pg::ScreenArgs sweep_screen_args(const pg_ctx* c) {
  pg::ScreenArgs a = screen_args(c);
  a.base16lo = nullptr;
  a.rnorms = c->rnorms;
  return a;
}

int prepare_screen(pg_ctx* c, const float* d_base) {
  PG_HIP(c, pg::launch_prepare_screen(d_base, c->p.n_base, c->p.dim, c->dpad,
                                      c->base16, c->base16lo, c->norms,
                                      c->rnorms, c->stream));
  if (c->base8)
    PG_HIP(c, pg::launch_prepare_fp8(d_base, c->p.n_base, c->p.dim, c->dpad,
                                     c->base8, c->scale8, c->stream));
  return PG_OK;
}

// Locality order of `n` query rows for the persistent search (tune_flags
// bit 9): *out = the order, or nullptr when the order is off, the grid is
// not persistent or n exceeds the order's buffers (n_base rows).  reuse:
// keep the order already built for the same rows pointer and count (the
// single-stage entry points; a build recomputes it).  Any permutation gives
// the same results, so an order built from rows that changed since only
// costs speed.
int search_order(pg_ctx* c, const float* d_rows, uint32_t n,
                 const uint32_t** out, bool reuse = false) {
  *out = nullptr;
  if (!c->order || !(c->p.tune_flags & kTunePersistent) || n == 0 ||
      n > c->p.n_base)
    return PG_OK;
  if (!(reuse && c->order_rows == d_rows && c->order_n == n)) {
    PG_HIP(c, pg::launch_locality_order(d_rows, n, c->p.dim, c->ogroup,
                                        c->ocnt, c->order, c->stream));
    c->order_rows = d_rows;
    c->order_n = n;
  }
  *out = c->order;
  return PG_OK;
}

// The whole stage on the context's stream.  prologue_done: the screened
// copies of the base and the match were already enqueued (the host entry
// point runs them chunk by chunk as the inputs arrive); proj_ready, when
// given, is recorded once the projection rows are final.
int run_pipeline(pg_ctx* c, const float* d_base, const uint32_t* d_gt,
                 uint32_t* d_proj, uint32_t* d_link,
                 bool prologue_done = false, hipEvent_t proj_ready = nullptr) {
  const pg_params& p = c->p;
  const uint32_t N = p.n_base;
  hipStream_t s = c->stream;
  int ev = 0;
  PG_TRY(record(c, ev++));
  const pg::ScreenArgs sc = screen_args(c);
  const pg::ScreenArgs* scp = c->base16 ? &sc : nullptr;
  // the link-row prune: the sweep screens plainly, the candidate-major
  // kernel with the context's screen
  const bool sweep = !(p.tune_flags & kTuneScanPrune);
  const pg::ScreenArgs scw = sweep_screen_args(c);
  const pg::ScreenArgs* scl = c->base16 ? (sweep ? &scw : &sc) : nullptr;
  if (!prologue_done) {
    if (c->base16)  // screened path: BF16 split of the base and row norms
      PG_TRY(prepare_screen(c, d_base));
    PG_HIP(c, pg::launch_match(d_gt, p.n_query, p.gt_k, N, c->winner, s));
  }
  PG_TRY(record(c, ev++));
  PG_HIP(c, pg::launch_prune(d_base, N, p.dim, nullptr, p.gt_k, d_gt,
                             c->winner, p.gt_k, p.proj_pr, c->pruned, c->pcnt,
                             c->stats, nullptr, false, nullptr, s));
  PG_TRY(record(c, ev++));
  PG_HIP(c, pg::launch_reverse(c->pruned, c->pcnt, N, p.proj_pr, c->deg,
                               c->off, c->cursor, c->scan_tmp, c->srcs, s));
  PG_TRY(record(c, ev++));
  PG_HIP(c, pg::launch_merge(d_base, N, p.dim, c->off, c->srcs, p.proj_rev,
                             c->pruned, c->pcnt, p.proj_pr, p.proj_deg, 0,
                             d_proj, c->stats, nullptr, nullptr, s));
  if (proj_ready) PG_HIP(c, hipEventRecord(proj_ready, s));
  PG_TRY(record(c, ev++));
  const uint32_t* graph = d_proj;
  uint32_t gdeg = p.proj_deg;
  const uint32_t* sorder = nullptr;
  PG_TRY(search_order(c, d_base, N, &sorder));
  // Screened paths: the search emits its first `rerank` entries (ordered by
  // screened distance) and the prune re-ranks exactly those in FP32.
  const uint32_t width =
      (c->base16 && p.rerank && p.rerank < p.topk) ? p.rerank : p.topk;
  for (uint32_t r = 0; r < p.rounds; ++r) {
    PG_HIP(c, pg::launch_search(d_base, N, d_base,
                                c->base8 ? nullptr : c->base16, c->base8,
                                c->scale8, c->dpad, N,
                                p.dim, graph, gdeg, p.Km,
                                p.Kp, p.Kd, width, p.reset_iter,
                                p.entry_seed, p.max_iters,
                                effective_table_size(p), tune_gather(p),
                                c->res, c->stats,
                                (p.tune_flags & kTunePersistent) ? c->queue
                                                                 : nullptr,
                                sorder, tune_waves(p), s));
    PG_TRY(record(c, ev++));
    PG_HIP(c, pg::launch_prune(d_base, N, p.dim, c->res, width, nullptr,
                               nullptr, 0, p.M, c->pruned, c->pcnt, c->stats,
                               scl, sweep, sorder, s));
    PG_TRY(record(c, ev++));
    PG_HIP(c, pg::launch_reverse(c->pruned, c->pcnt, N, p.M, c->deg, c->off,
                                 c->cursor, c->scan_tmp, c->srcs, s));
    PG_TRY(record(c, ev++));
    PG_HIP(c, pg::launch_merge(d_base, N, p.dim, c->off, c->srcs, p.link_rev,
                               c->pruned, c->pcnt, p.M, p.M, 1, d_link,
                               c->stats, scp,
                               (p.tune_flags & kTuneOrderMerge) ? sorder
                                                                : nullptr,
                               s));
    PG_TRY(record(c, ev++));
    graph = d_link;
    gdeg = p.M;
  }
  c->n_stage_events = c->capturing ? 0 : ev;
  if (!c->capturing) PG_HIP(c, hipEventRecord(c->done, s));
  return PG_OK;
}

void drop_graph(pg_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

// Graph mode: the pipeline (kernels only: every clear is a fill kernel, no
// memset node) is captured once per set of buffers and replayed.  Stage
// events are not recorded in graph mode, so per-stage times are reported by
// eager builds only.
int run_pipeline_graph(pg_ctx* c, const float* d_base, const uint32_t* d_gt,
                       uint32_t* d_proj, uint32_t* d_link) {
  const void* key[4] = {d_base, d_gt, d_proj, d_link};
  if (!c->graph_exec || std::memcmp(key, c->graph_key, sizeof(key)) != 0) {
    drop_graph(c);
    PG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    const int rc = run_pipeline(c, d_base, d_gt, d_proj, d_link);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != PG_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return set_err(c, PG_ERR_HIP, "hipStreamEndCapture: %s",
                     hipGetErrorString(e));
    c->graph = g;
    PG_HIP(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr,
                                  0));
    std::memcpy(c->graph_key, key, sizeof(key));
  }
  PG_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
  PG_HIP(c, hipEventRecord(c->done, c->stream));
  c->n_stage_events = 0;
  return PG_OK;
}

}  // namespace

extern "C" {

int pg_params_init(pg_params* p, uint32_t n_base, uint32_t dim,
                   uint32_t n_query, uint32_t M, uint32_t L) {
  if (!p) return PG_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->n_base = n_base;
  p->dim = dim;
  p->n_query = n_query;
  p->gt_k = 128;
  p->proj_deg = 16;
  p->proj_pr = 15;
  p->proj_rev = 111;
  p->M = M;
  p->link_rev = M < 126 ? 126 - M : 0;
  p->Km = L;
  p->Kd = 16;
  p->Kp = (L + 15) / 16;
  p->topk = 2 * L;
  p->reset_iter = 15;
  p->table_size = 0;  // auto: see effective_table_size()
  p->rounds = 2;
  p->entry_seed = 3030517;
  p->max_iters = 100000;
  p->precision = PG_FP32;
  p->rerank = 0;
  p->tune_flags = kTunePersistent;  // measured faster on MI355X
  return PG_OK;
}

int pg_params_check(const pg_params* p, char* msg, uint32_t msg_len) {
  char buf[256] = "";
  auto fail = [&](const char* why) {
    std::snprintf(buf, sizeof(buf), "%s", why);
    if (msg && msg_len) std::snprintf(msg, msg_len, "%s", buf);
    return PG_ERR_INVALID;
  };
  if (!p) return fail("params is NULL");
  if (p->n_base == 0 || p->n_base >= 0x80000000u)
    return fail("n_base must be in [1, 2^31) (the id MSB is the expanded flag)");
  if (p->dim == 0 || p->dim > 1024) return fail("dim must be in [1, 1024]");
  if (p->gt_k < 3 || p->gt_k > 512) return fail("gt_k must be in [3, 512]");
  if (p->proj_pr < 2 || p->proj_pr > 128) return fail("proj_pr must be in [2, 128]");
  if (p->proj_deg < 1 || p->proj_deg > 128) return fail("proj_deg must be in [1, 128]");
  if (p->proj_rev + p->proj_pr > 128)
    return fail("proj_rev + proj_pr must be <= 128 (merge list width)");
  if (p->M < 2 || p->M > 128) return fail("M must be in [2, 128]");
  if (p->link_rev + p->M > 128) return fail("link_rev + M must be <= 128");
  if (p->Km == 0 || p->Kp == 0 || p->Kd == 0) return fail("Km, Kp, Kd must be > 0");
  if (p->Kd > p->proj_deg || (p->rounds > 1 && p->Kd > p->M))
    return fail("Kd must not exceed the degree of the graph it scans");
  if (p->topk == 0 || p->topk > p->Km + p->Kp * p->Kd)
    return fail("topk must be in [1, Km + Kp*Kd]");
  if (p->topk > 512) return fail("topk must be <= 512");
  if (p->Km + p->Kp * p->Kd > 4096) return fail("Km + Kp*Kd must be <= 4096");
  if (p->reset_iter == 0) return fail("reset_iter must be > 0");
  if (p->Kp * p->Kd < p->Km)
    return fail("Kp*Kd must be >= Km (the entry points fill the first list)");
  if (p->precision != PG_FP32) {
    if (4ull * (p->Km + uint64_t(p->Kp) * p->Kd) >
        3ull * effective_table_size(*p) + 3)
      return fail("screened path: table_size must be >= 4/3 (Km + Kp*Kd)");
  } else if (uint64_t(effective_table_size(*p)) <=
             uint64_t(p->Km) + uint64_t(p->reset_iter) * p->Kp * p->Kd)
    return fail("table_size must exceed Km + reset_iter*Kp*Kd (else the "
                "visited table can fill up)");
  {
    uint32_t lp2 = 64;
    while (lp2 < p->Km + p->Kp * p->Kd) lp2 <<= 1;
    if (pg::search_smem_bytes(effective_table_size(*p), lp2, p->Kp * p->Kd,
                              p->Kp) >
        160u * 1024u)
      return fail("search LDS footprint exceeds 160 KiB");
  }
  if (p->rounds > 3) return fail("rounds must be <= 3");
  if (p->max_iters == 0) return fail("max_iters must be > 0");
  if (p->rerank > p->topk) return fail("rerank must be <= topk (0 = topk)");
  if (p->precision != PG_FP32 && p->precision != PG_BF16 &&
      p->precision != PG_FP8E4M3)
    return fail("precision must be PG_FP32, PG_BF16 or PG_FP8E4M3");
  if (p->precision != PG_FP32 && p->dim > 512)
    return fail("the screened paths support dim <= 512");

  if (msg && msg_len) msg[0] = 0;
  return PG_OK;
}

pg_ctx* pg_create(int device, const pg_params* p) {
  char msg[256];
  if (pg_params_check(p, msg, sizeof(msg)) != PG_OK) {
    set_err(nullptr, PG_ERR_INVALID, "invalid params: %s", msg);
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err(nullptr, PG_ERR_NO_DEVICE, "no HIP device available");
    return nullptr;
  }
  if (device < 0 || device >= ndev) {
    set_err(nullptr, PG_ERR_INVALID, "device %d out of range (%d devices)",
            device, ndev);
    return nullptr;
  }
  pg_ctx* c = new (std::nothrow) pg_ctx();
  if (!c) {
    set_err(nullptr, PG_ERR_OOM, "host allocation failed");
    return nullptr;
  }
  c->device = device;
  c->p = *p;
  c->mp_max = p->proj_pr > p->M ? p->proj_pr : p->M;
  c->dpad = (p->dim + 31) / 32 * 32;
  c->eps_c = (p->tune_flags & kTuneSplit) ? screen_eps_split(p->dim, c->dpad)
                                           : screen_eps(p->dim, c->dpad);
  // plain screen, per-pair bound with residual norms: gamma factors of the
  // MFMA accumulation over dpad terms (2^-23, any rounding of the adds) and
  // of the exact chain over dim terms (2^-24)
  {
    const double ua = std::ldexp(1.0, -23), uc = std::ldexp(1.0, -24);
    c->gam_a = float(c->dpad * ua / (1.0 - c->dpad * ua));
    c->gam_c = float(p->dim * uc / (1.0 - p->dim * uc));
  }
  auto fail = [&](int) {
    std::memcpy(g_err, c->err, sizeof(g_err));
    pg_destroy(c);
    return static_cast<pg_ctx*>(nullptr);
  };
  if (hipSetDevice(device) != hipSuccess) {
    set_err(c, PG_ERR_HIP, "hipSetDevice(%d) failed", device);
    return fail(0);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->done) != hipSuccess) {
    set_err(c, PG_ERR_HIP, "stream/event creation failed");
    return fail(0);
  }
  for (int i = 0; i <= kMaxStages; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) {
      set_err(c, PG_ERR_HIP, "event creation failed");
      return fail(0);
    }
  const uint64_t N = p->n_base;
  const uint32_t nb = (p->n_base + 1023) / 1024 + 1;
  if (dev_alloc(c, &c->winner, N, "winner") ||
      dev_alloc(c, &c->pruned, N * c->mp_max, "pruned") ||
      dev_alloc(c, &c->pcnt, N, "pcnt") || dev_alloc(c, &c->deg, N, "deg") ||
      dev_alloc(c, &c->cursor, N, "cursor") ||
      dev_alloc(c, &c->off, N + 1, "off") ||
      dev_alloc(c, &c->scan_tmp, nb, "scan_tmp") ||
      dev_alloc(c, &c->srcs, N * c->mp_max, "srcs") ||
      dev_alloc(c, &c->res, N * p->topk, "search_out") ||
      (p->precision != PG_FP32 &&
       (dev_alloc(c, reinterpret_cast<uint16_t**>(&c->base16),
                  N * ((p->dim + 31) / 32 * 32), "base_bf16") ||
        dev_alloc(c, &c->norms, N, "norms") ||
        dev_alloc(c, &c->rnorms, N, "rnorms"))) ||
      (p->precision != PG_FP32 && (p->tune_flags & kTuneSplit) &&
       dev_alloc(c, reinterpret_cast<uint16_t**>(&c->base16lo),
                 N * ((p->dim + 31) / 32 * 32), "base_bf16_lo")) ||
      (p->precision == PG_FP8E4M3 &&
       (dev_alloc(c, reinterpret_cast<uint8_t**>(&c->base8),
                  N * ((p->dim + 31) / 32 * 32), "base_fp8") ||
        dev_alloc(c, &c->scale8, N, "scale8"))) ||
      dev_alloc(c, &c->stats, kStatCount, "stats") ||
      dev_alloc(c, &c->stage_cnt, (kMaxStages + 1) * 16ull, "stage_cnt") ||
      dev_alloc(c, &c->queue, pg::kQueueSegs, "queue") ||
      ((p->tune_flags & kTuneOrder) &&
       (dev_alloc(c, &c->order, N, "order") ||
        dev_alloc(c, &c->ogroup, N, "order_group") ||
        dev_alloc(c, &c->ocnt, 2ull * pg::kOrderGroups, "order_cnt"))))
    return fail(0);
  if (pg::launch_fill_u64(c->stats, kStatCount, 0ull, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    set_err(c, PG_ERR_HIP, "initial clear failed");
    return fail(0);
  }
  return c;
}

void pg_destroy(pg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->aux) {
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
    c->aux = nullptr;
  }
  if (c->copy) {
    (void)hipStreamSynchronize(c->copy);
    (void)hipStreamDestroy(c->copy);
    c->copy = nullptr;
  }
  for (hipEvent_t& e : c->ev_chunk)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_proj) (void)hipEventDestroy(c->ev_proj);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_bip) (void)hipEventDestroy(c->ev_bip);
  c->ev_in = c->ev_bip = nullptr;
  dev_free(c->winner);
  dev_free(c->pruned);
  dev_free(c->pcnt);
  dev_free(c->deg);
  dev_free(c->cursor);
  dev_free(c->off);
  dev_free(c->scan_tmp);
  dev_free(c->srcs);
  dev_free(c->res);
  if (c->base16) (void)hipFree(c->base16);
  if (c->base8) (void)hipFree(c->base8);
  if (c->base16lo) (void)hipFree(c->base16lo);
  c->base16 = c->base8 = c->base16lo = nullptr;
  dev_free(c->scale8);
  dev_free(c->norms);
  dev_free(c->rnorms);
  drop_graph(c);
  dev_free(c->stats);
  dev_free(c->stage_cnt);
  dev_free(c->queue);
  dev_free(c->order);
  dev_free(c->ogroup);
  dev_free(c->ocnt);
  dev_free(c->fuse_ws);
  dev_free(c->h_base_dev);
  dev_free(c->h_gt_dev);
  dev_free(c->h_proj_dev);
  dev_free(c->h_link_dev);
  for (int i = 0; i <= kMaxStages; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->done) (void)hipEventDestroy(c->done);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int pg_gpu_stage_device_async(pg_ctx* c, const float* d_base,
                              const uint32_t* d_gt, uint32_t* d_proj,
                              uint32_t* d_link) {
  if (!c) return set_err(nullptr, PG_ERR_INVALID, "ctx is NULL");
  if (!d_base || !d_gt || !d_proj || !d_link)
    return set_err(c, PG_ERR_INVALID, "NULL buffer");
  PG_HIP(c, hipSetDevice(c->device));
  if (c->p.tune_flags & kTuneGraph)
    return run_pipeline_graph(c, d_base, d_gt, d_proj, d_link);
  return run_pipeline(c, d_base, d_gt, d_proj, d_link);
}

int pg_gpu_stage_async(pg_ctx* c, const float* base, const uint32_t* gt,
                       uint32_t* out_proj, uint32_t* out_link) {
  if (!c) return set_err(nullptr, PG_ERR_INVALID, "ctx is NULL");
  if (!base || !gt || !out_proj || !out_link)
    return set_err(c, PG_ERR_INVALID, "NULL buffer");
  const pg_params& p = c->p;
  const uint64_t N = p.n_base;
  PG_HIP(c, hipSetDevice(c->device));
  if (!c->h_base_dev) {
    PG_TRY(dev_alloc(c, &c->h_base_dev, N * p.dim, "base"));
    PG_TRY(dev_alloc(c, &c->h_gt_dev, uint64_t(p.n_query) * p.gt_k, "gt"));
    PG_TRY(dev_alloc(c, &c->h_proj_dev, N * p.proj_deg, "proj"));
    PG_TRY(dev_alloc(c, &c->h_link_dev, N * p.M, "link"));
  }
  if (c->p.tune_flags & kTuneGraph) {
    // the captured pipeline is replayed whole: plain copies around it
    PG_HIP(c, hipMemcpyAsync(c->h_base_dev, base, N * p.dim * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
    PG_HIP(c, hipMemcpyAsync(c->h_gt_dev, gt,
                             uint64_t(p.n_query) * p.gt_k * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->stream));
    PG_TRY(run_pipeline_graph(c, c->h_base_dev, c->h_gt_dev, c->h_proj_dev,
                              c->h_link_dev));
    PG_HIP(c, hipMemcpyAsync(out_proj, c->h_proj_dev,
                             N * p.proj_deg * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
    PG_HIP(c, hipMemcpyAsync(out_link, c->h_link_dev,
                             N * p.M * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
    PG_HIP(c, hipEventRecord(c->done, c->stream));
    return PG_OK;
  }
  if (!c->copy) {
    PG_HIP(c, hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking));
    for (hipEvent_t& e : c->ev_chunk)
      PG_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    PG_HIP(c, hipEventCreateWithFlags(&c->ev_proj, hipEventDisableTiming));
    PG_HIP(c, hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
  }
  // Inputs stream in on the copy stream in about 8 chunks each (256 KiB to
  // 64 MiB per chunk): the gt first (the match runs on each chunk as it
  // lands), then the base (the screened copies are made per chunk).  Copies from pinned host memory
  // (pg_host_alloc) run at full link speed; pageable memory also works,
  // staged by the runtime.  The copy stream first waits for the main
  // stream, so a previous build's kernels are done with the buffers.
  hipStream_t s = c->stream;
  PG_HIP(c, hipEventRecord(c->ev_out, s));
  PG_HIP(c, hipStreamWaitEvent(c->copy, c->ev_out, 0));
  PG_HIP(c, pg::launch_fill_u32(c->winner, N, PG_TOMB, s));
  uint32_t next_ev = 0;
  auto chunked = [&](uint64_t rows, uint64_t row_bytes, auto&& copy_rows,
                     auto&& on_rows) -> int {
    uint64_t chunk = rows * row_bytes / 8;
    chunk = chunk < (256ull << 10) ? (256ull << 10)
            : chunk > (64ull << 20) ? (64ull << 20) : chunk;
    const uint64_t per = row_bytes >= chunk ? 1 : chunk / row_bytes;
    for (uint64_t r0 = 0; r0 < rows; r0 += per) {
      const uint64_t nr = rows - r0 < per ? rows - r0 : per;
      PG_TRY(copy_rows(r0, nr));
      hipEvent_t e = c->ev_chunk[next_ev++ % 8];
      PG_HIP(c, hipEventRecord(e, c->copy));
      PG_HIP(c, hipStreamWaitEvent(s, e, 0));
      PG_TRY(on_rows(r0, nr));
    }
    return PG_OK;
  };
  const uint64_t gt_row = uint64_t(p.gt_k) * sizeof(uint32_t);
  PG_TRY(chunked(
      p.n_query, gt_row,
      [&](uint64_t r0, uint64_t nr) -> int {
        PG_HIP(c, hipMemcpyAsync(c->h_gt_dev + r0 * p.gt_k, gt + r0 * p.gt_k,
                                 nr * gt_row, hipMemcpyHostToDevice, c->copy));
        return PG_OK;
      },
      [&](uint64_t r0, uint64_t nr) -> int {
        PG_HIP(c, pg::launch_match_rows(c->h_gt_dev + r0 * p.gt_k,
                                        uint32_t(r0), uint32_t(nr), p.gt_k,
                                        p.n_base, c->winner, s));
        return PG_OK;
      }));
  const uint64_t base_row = uint64_t(p.dim) * sizeof(float);
  PG_TRY(chunked(
      N, base_row,
      [&](uint64_t r0, uint64_t nr) -> int {
        PG_HIP(c, hipMemcpyAsync(c->h_base_dev + r0 * p.dim, base + r0 * p.dim,
                                 nr * base_row, hipMemcpyHostToDevice,
                                 c->copy));
        return PG_OK;
      },
      [&](uint64_t r0, uint64_t nr) -> int {
        if (!c->base16) return PG_OK;
        const uint64_t d = c->dpad;
        PG_HIP(c, pg::launch_prepare_screen(
                      c->h_base_dev + r0 * p.dim, uint32_t(nr), p.dim, c->dpad,
                      static_cast<uint16_t*>(c->base16) + r0 * d,
                      c->base16lo ? static_cast<uint16_t*>(c->base16lo) + r0 * d
                                  : nullptr,
                      c->norms + r0, c->rnorms + r0, s));
        if (c->base8)
          PG_HIP(c, pg::launch_prepare_fp8(
                        c->h_base_dev + r0 * p.dim, uint32_t(nr), p.dim,
                        c->dpad, static_cast<uint8_t*>(c->base8) + r0 * d,
                        c->scale8 + r0, s));
        return PG_OK;
      }));
  PG_TRY(run_pipeline(c, c->h_base_dev, c->h_gt_dev, c->h_proj_dev,
                      c->h_link_dev, /*prologue_done=*/true, c->ev_proj));
  // the projection rows leave while the link rounds run
  PG_HIP(c, hipStreamWaitEvent(c->copy, c->ev_proj, 0));
  PG_HIP(c, hipMemcpyAsync(out_proj, c->h_proj_dev,
                           N * p.proj_deg * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, c->copy));
  PG_HIP(c, hipMemcpyAsync(out_link, c->h_link_dev, N * p.M * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
  // the main stream's completion covers the copy stream too
  PG_HIP(c, hipEventRecord(c->ev_out, c->copy));
  PG_HIP(c, hipStreamWaitEvent(s, c->ev_out, 0));
  PG_HIP(c, hipEventRecord(c->done, s));
  return PG_OK;
}

int pg_host_alloc(uint64_t bytes, void** out) {
  if (!out) return set_err(nullptr, PG_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return set_err(nullptr, PG_ERR_INVALID, "zero bytes");
  const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return set_err(nullptr, PG_ERR_OOM, "hipHostMalloc(%llu bytes): %s",
                   (unsigned long long)bytes, hipGetErrorString(e));
  }
  return PG_OK;
}

int pg_host_free(void* ptr) {
  if (!ptr) return PG_OK;
  const hipError_t e = hipHostFree(ptr);
  if (e != hipSuccess)
    return set_err(nullptr, PG_ERR_HIP, "hipHostFree: %s",
                   hipGetErrorString(e));
  return PG_OK;
}

int pg_wait(pg_ctx* c) {
  if (!c) return set_err(nullptr, PG_ERR_INVALID, "ctx is NULL");
  PG_HIP(c, hipStreamSynchronize(c->stream));
  if (c->aux) PG_HIP(c, hipStreamSynchronize(c->aux));
  PG_HIP(c, hipGetLastError());
  return PG_OK;
}

int pg_sync(pg_ctx* c) { return pg_wait(c); }

int pg_query(pg_ctx* c) {
  if (!c) return set_err(nullptr, PG_ERR_INVALID, "ctx is NULL");
  hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess && c->aux) e = hipStreamQuery(c->aux);
  if (e == hipSuccess) return PG_OK;
  if (e == hipErrorNotReady) return PG_ERR_BUSY;
  return set_err(c, PG_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int pg_reset_stats(pg_ctx* c) {
  if (!c) return set_err(nullptr, PG_ERR_INVALID, "ctx is NULL");
  PG_HIP(c, pg::launch_fill_u64(c->stats, kStatCount, 0ull, c->stream));
  return pg_wait(c);
}

int pg_get_stats(pg_ctx* c, pg_stats* out) {
  if (!c || !out) return set_err(c, PG_ERR_INVALID, "NULL argument");
  PG_TRY(pg_wait(c));
  static_assert(kStatCount == 4096, "stat layout");
  unsigned long long host[kStatCount];
  PG_HIP(c, hipMemcpy(host, c->stats, sizeof(host), hipMemcpyDeviceToHost));
  uint64_t f[16] = {};
  for (int s = 0; s < 256; ++s)
    for (int k = 0; k < 16; ++k) f[k] += host[s * 16 + k];
  std::memset(out, 0, sizeof(*out));
  out->prune_dist = f[0];
  out->prune_pairs = f[1];
  out->merge_dist = f[2];
  out->merge_pairs = f[3];
  out->search_dist = f[4];
  out->search_iters = f[5];
  out->search_capped = f[6];
  out->rev_overflow = f[7];
  out->screen_fallback = f[8];
  out->n_stages = c->n_stage_events > 1 ? uint32_t(c->n_stage_events - 1) : 0u;
  for (uint32_t i = 0; i < out->n_stages && i < 16; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]) == hipSuccess)
      out->stage_ms[i] = ms;
  }
  return PG_OK;
}

int pg_get_stage_counters(pg_ctx* c, uint32_t stage, uint64_t* out16) {
  if (!c || !out16) return set_err(c, PG_ERR_INVALID, "NULL argument");
  if (c->n_stage_events < 2 || stage + 1 >= uint32_t(c->n_stage_events))
    return set_err(c, PG_ERR_INVALID, "stage %u not recorded", stage);
  PG_TRY(pg_wait(c));
  unsigned long long rows[2][16];
  PG_HIP(c, hipMemcpy(rows, c->stage_cnt + stage * 16, sizeof(rows),
                      hipMemcpyDeviceToHost));
  for (int k = 0; k < 16; ++k) out16[k] = rows[1][k] - rows[0][k];
  return PG_OK;
}

const char* pg_last_error(const pg_ctx* c) { return c ? c->err : g_err; }

const char* pg_stage_name(uint32_t i) {
  return i < uint32_t(kMaxStages) ? kStageNames[i] : "";
}

int pg_run_match(pg_ctx* c, const uint32_t* d_gt, uint32_t n_query,
                 uint32_t gt_k, uint32_t* d_winner) {
  if (!c || !d_winner || (n_query && !d_gt))
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  PG_HIP(c, hipSetDevice(c->device));
  PG_HIP(c, pg::launch_match(d_gt, n_query, gt_k, c->p.n_base, d_winner,
                             c->stream));
  return PG_OK;
}

int pg_run_prune(pg_ctx* c, const float* d_base, const uint32_t* d_rows,
                 uint32_t D, const uint32_t* d_gt, const uint32_t* d_winner,
                 uint32_t gt_k, uint32_t mp, uint32_t* d_pruned,
                 uint32_t* d_pcnt) {
  if (!c || !d_base || !d_pruned || !d_pcnt)
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  if (!d_rows) {
    if (!d_gt || !d_winner) return set_err(c, PG_ERR_INVALID, "top-1 mode needs gt and winner");
    if (gt_k < 3 || gt_k > 512) return set_err(c, PG_ERR_INVALID, "gt_k out of range");
    D = gt_k;
  }
  if (D == 0 || D > 512) return set_err(c, PG_ERR_INVALID, "D must be in [1, 512]");
  if (mp < 2 || mp > 128) return set_err(c, PG_ERR_INVALID, "mp must be in [2, 128]");
  PG_HIP(c, hipSetDevice(c->device));
  const bool sweep = !(c->p.tune_flags & kTuneScanPrune);
  // the sweep screens plainly (link rows), the candidate-major kernel with
  // the context's screen
  const pg::ScreenArgs sc =
      (sweep && d_rows) ? sweep_screen_args(c) : screen_args(c);
  if (c->base16) PG_TRY(prepare_screen(c, d_base));
  // link rows take nodes in the locality order when the context has one
  const uint32_t* sorder = nullptr;
  if (d_rows) PG_TRY(search_order(c, d_base, c->p.n_base, &sorder, true));
  PG_HIP(c, pg::launch_prune(d_base, c->p.n_base, c->p.dim, d_rows, D, d_gt,
                             d_winner, gt_k, mp, d_pruned, d_pcnt, c->stats,
                             c->base16 ? &sc : nullptr, sweep, sorder,
                             c->stream));
  return PG_OK;
}

int pg_run_reverse(pg_ctx* c, const uint32_t* d_pruned, const uint32_t* d_pcnt,
                   uint32_t mp, uint32_t* d_off, uint32_t* d_srcs) {
  if (!c || !d_pruned || !d_pcnt || !d_off || !d_srcs)
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  if (mp == 0 || mp > 128) return set_err(c, PG_ERR_INVALID, "mp must be in [1, 128]");
  PG_HIP(c, hipSetDevice(c->device));
  PG_HIP(c, pg::launch_reverse(d_pruned, d_pcnt, c->p.n_base, mp, c->deg, d_off,
                               c->cursor, c->scan_tmp, d_srcs, c->stream));
  return PG_OK;
}

int pg_run_merge(pg_ctx* c, const float* d_base, const uint32_t* d_off,
                 const uint32_t* d_srcs, uint32_t cap,
                 const uint32_t* d_pruned, const uint32_t* d_pcnt, uint32_t mp,
                 uint32_t deg, int strict, uint32_t* d_out) {
  if (!c || !d_base || !d_off || !d_srcs || !d_pruned || !d_pcnt || !d_out)
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  if (cap + mp > 128) return set_err(c, PG_ERR_INVALID, "cap + mp must be <= 128");
  if (deg == 0 || deg > 128) return set_err(c, PG_ERR_INVALID, "deg must be in [1, 128]");
  PG_HIP(c, hipSetDevice(c->device));
  const pg::ScreenArgs sc = screen_args(c);
  if (c->base16) PG_TRY(prepare_screen(c, d_base));
  const uint32_t* sorder = nullptr;
  if (c->p.tune_flags & kTuneOrderMerge)
    PG_TRY(search_order(c, d_base, c->p.n_base, &sorder, true));
  PG_HIP(c, pg::launch_merge(d_base, c->p.n_base, c->p.dim, d_off, d_srcs, cap,
                             d_pruned, d_pcnt, mp, deg, strict, d_out,
                             c->stats, c->base16 ? &sc : nullptr, sorder,
                             c->stream));
  return PG_OK;
}

int pg_run_search(pg_ctx* c, const float* d_queries, uint32_t n_query,
                  const float* d_base, const uint32_t* d_graph,
                  uint32_t graph_deg, uint32_t* d_out) {
  if (!c || !d_base || !d_graph || !d_out || (n_query && !d_queries))
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  const pg_params& p = c->p;
  if (p.Kd > graph_deg)
    return set_err(c, PG_ERR_INVALID, "Kd (%u) exceeds graph degree (%u)", p.Kd,
                   graph_deg);
  PG_HIP(c, hipSetDevice(c->device));
  if (c->base16) PG_TRY(prepare_screen(c, d_base));
  const uint32_t* sorder = nullptr;
  PG_TRY(search_order(c, d_queries, n_query, &sorder, true));
  PG_HIP(c, pg::launch_search(d_queries, n_query, d_base,
                              c->base8 ? nullptr : c->base16, c->base8,
                              c->scale8, c->dpad,
                              p.n_base, p.dim,
                              d_graph, graph_deg, p.Km, p.Kp, p.Kd, p.topk,
                              p.reset_iter, p.entry_seed, p.max_iters,
                              effective_table_size(p), tune_gather(p), d_out,
                              c->stats,
                              (p.tune_flags & kTunePersistent) ? c->queue
                                                               : nullptr,
                              sorder, tune_waves(p), c->stream));
  return PG_OK;
}

int pg_run_pair_distances(pg_ctx* c, const float* d_queries,
                          uint32_t n_query, const float* d_base,
                          const uint32_t* d_ids, uint32_t k, int kind,
                          float* d_out) {
  if (!c || !d_base || (n_query && (!d_queries || !d_ids || !d_out)))
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  if (k > 1024) return set_err(c, PG_ERR_INVALID, "k (%u) exceeds 1024", k);
  if (kind != 0 && kind != 1)
    return set_err(c, PG_ERR_INVALID, "kind must be 0 (search) or 1 (chain)");
  const pg_params& p = c->p;
  PG_HIP(c, hipSetDevice(c->device));
  if (kind == 0 && c->base16) PG_TRY(prepare_screen(c, d_base));
  PG_HIP(c, pg::launch_pair_dist(d_queries, n_query, d_base,
                                 c->base8 ? nullptr : c->base16, c->base8,
                                 c->scale8, c->dpad, p.n_base, p.dim, d_ids,
                                 k, kind, d_out, c->stream));
  return PG_OK;
}

}  // extern "C"

namespace {

int check_fusion_shape(pg_ctx* c, const pg_host_params* hp) {
  const pg_params& p = c->p;
  if (hp->n_base != p.n_base || hp->dim != p.dim)
    return set_err(c, PG_ERR_INVALID,
                   "host params must have the context's n_base and dim");
  if (hp->n_query != hp->n_base)
    return set_err(c, PG_ERR_INVALID, "the GPU host stages need n_query == n_base");
  if (hp->gt_k == 0 || hp->gt_k > 256)
    return set_err(c, PG_ERR_INVALID, "gt_k must be in [1, 256]");
  if (hp->M_nn == 0 || hp->M_nn > 64)
    return set_err(c, PG_ERR_INVALID, "M_nn must be in [1, 64]");
  if (hp->rounds >= hp->gt_k || uint64_t(hp->rounds) * hp->M_nn > 256)
    return set_err(c, PG_ERR_INVALID,
                   "rounds must be < gt_k and rounds * M_nn <= 256");
  if (hp->M_final == 0 || hp->M_final > 128)
    return set_err(c, PG_ERR_INVALID, "M_final must be in [1, 128]");
  return PG_OK;
}

pg::FusionShape fusion_shape(const pg_host_params* hp, uint32_t M_link,
                             uint32_t proj_deg) {
  return pg::FusionShape{hp->n_base, hp->dim, hp->gt_k, hp->M_nn,
                         hp->rounds, hp->M_final, M_link, proj_deg};
}

// Workspace for shape f (kept across calls; grown after both streams are
// idle).
int fusion_workspace(pg_ctx* c, const pg::FusionShape& f) {
  const uint64_t words = pg::fusion_ws_words(f);
  if (c->fuse_words >= words) return PG_OK;
  PG_HIP(c, hipStreamSynchronize(c->stream));
  if (c->aux) PG_HIP(c, hipStreamSynchronize(c->aux));
  dev_free(c->fuse_ws);
  c->fuse_words = 0;
  c->bip_valid = false;
  PG_TRY(dev_alloc(c, &c->fuse_ws, words, "fusion workspace"));
  c->fuse_words = words;
  return PG_OK;
}

bool same_bip_shape(const pg_host_params& a, const pg_host_params& b) {
  return a.n_base == b.n_base && a.dim == b.dim && a.n_query == b.n_query &&
         a.gt_k == b.gt_k && a.M_nn == b.M_nn && a.rounds == b.rounds &&
         a.M_final == b.M_final;
}

}  // namespace

extern "C" {

int pg_run_bipartite_async(pg_ctx* c, const pg_host_params* hp,
                           const float* d_base, const uint32_t* d_gt) {
  if (!c || !hp || !d_base || !d_gt)
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  PG_TRY(check_fusion_shape(c, hp));
  PG_HIP(c, hipSetDevice(c->device));
  if (!c->aux) {
    PG_HIP(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    PG_HIP(c, hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    PG_HIP(c, hipEventCreateWithFlags(&c->ev_bip, hipEventDisableTiming));
  }
  // the workspace is sized for the widest FusionFinal candidate rows
  const pg::FusionShape f = fusion_shape(hp, 0, 0);
  PG_TRY(fusion_workspace(c, f));
  // after everything already on the context's stream, on the second stream
  PG_HIP(c, hipEventRecord(c->ev_in, c->stream));
  PG_HIP(c, hipStreamWaitEvent(c->aux, c->ev_in, 0));
  PG_HIP(c, pg::launch_fusion_bipartite(f, d_base, d_gt, c->fuse_ws, c->aux));
  PG_HIP(c, hipEventRecord(c->ev_bip, c->aux));
  c->bip_valid = true;
  c->bip_hp = *hp;
  return PG_OK;
}

int pg_run_fusion_final(pg_ctx* c, const pg_host_params* hp,
                        const float* d_base, const uint32_t* d_proj,
                        uint32_t proj_deg, const uint32_t* d_link,
                        uint32_t M_link, uint32_t* d_off, uint32_t* d_ids) {
  if (!c || !hp || !d_base || !d_off || !d_ids || (M_link && !d_link) ||
      (proj_deg && !d_proj))
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  PG_TRY(check_fusion_shape(c, hp));
  if (uint64_t(2) * hp->M_nn + M_link + proj_deg > 256)
    return set_err(c, PG_ERR_INVALID,
                   "2 * M_nn + M_link + proj_deg must be <= 256");
  if (!c->bip_valid || !same_bip_shape(c->bip_hp, *hp))
    return set_err(c, PG_ERR_INVALID,
                   "no bipartite rows of this shape: call "
                   "pg_run_bipartite_async (or pg_run_fusion) first");
  PG_HIP(c, hipSetDevice(c->device));
  if (c->aux) PG_HIP(c, hipStreamWaitEvent(c->stream, c->ev_bip, 0));
  PG_HIP(c, pg::launch_fusion_final(fusion_shape(hp, M_link, proj_deg),
                                    d_base, d_proj, d_link, c->fuse_ws, d_off,
                                    d_ids, c->stream));
  return PG_OK;
}

int pg_run_fusion(pg_ctx* c, const pg_host_params* hp, const float* d_base,
                  const uint32_t* d_gt, const uint32_t* d_proj,
                  uint32_t proj_deg, const uint32_t* d_link, uint32_t M_link,
                  uint32_t* d_off, uint32_t* d_ids) {
  if (!c || !hp || !d_base || !d_gt || !d_off || !d_ids ||
      (M_link && !d_link) || (proj_deg && !d_proj))
    return set_err(c, PG_ERR_INVALID, "NULL argument");
  PG_TRY(check_fusion_shape(c, hp));
  if (uint64_t(2) * hp->M_nn + M_link + proj_deg > 256)
    return set_err(c, PG_ERR_INVALID,
                   "2 * M_nn + M_link + proj_deg must be <= 256");
  PG_HIP(c, hipSetDevice(c->device));
  const pg::FusionShape f = fusion_shape(hp, M_link, proj_deg);
  PG_TRY(fusion_workspace(c, f));
  // a bipartite stage still running on the second stream shares the
  // workspace: order after it
  if (c->aux) PG_HIP(c, hipStreamWaitEvent(c->stream, c->ev_bip, 0));
  PG_HIP(c, pg::launch_fusion(f, d_base, d_gt, d_proj, d_link, c->fuse_ws,
                              d_off, d_ids, c->stream));
  c->bip_valid = true;
  c->bip_hp = *hp;
  return PG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free search (pg_search_noalloc).  The reference has no
// counterpart; it serves callers that cannot allocate device memory (code
// under hipGraph capture, serving loops).  Its device scratch is one
// file-scope array per GPU, held until the process ends:
//   words [0, 64)                     persistent-search queue counters
//   words [64, 64 + 2*kOrderGroups)   locality-order group counters
//   then kNoAllocMaxQueries words     pivot group of each query
//   then kNoAllocMaxQueries words     the locality order
// Host code never uses the array's name as an address: the device address
// comes from hipGetSymbolAddress, once per device, on the first call made
// with that device current.  Every call on one GPU shares the array, so
// calls that overlap on one GPU race; every region a call reads is
// rewritten or cleared (by fill kernels, no memset) inside that call.
namespace {

constexpr uint32_t kNoAllocMaxQueries = 1u << 20;
constexpr size_t kScratchQueueWords = 64;
constexpr size_t kScratchCntWords = 2ull * pg::kOrderGroups;
constexpr size_t kScratchWords = kScratchQueueWords + kScratchCntWords +
                                 2ull * kNoAllocMaxQueries;
static_assert(pg::kQueueSegs <= kScratchQueueWords, "queue counters fit");

static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_address(int dev, uint32_t** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(nullptr, PG_ERR_INVALID,
                   "pg_search_noalloc: device %d beyond the %d supported", dev,
                   kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return set_err(nullptr, PG_ERR_HIP,
                     "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                     hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<uint32_t*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return PG_OK;
}

}  // namespace

extern "C" {

int pg_search_noalloc(const pg_params* p, const float* d_queries,
                      uint32_t n_query, const float* d_base,
                      const uint32_t* d_graph, uint32_t graph_deg,
                      uint32_t* d_out, void* stream) {
  char msg[256];
  if (pg_params_check(p, msg, sizeof(msg)) != PG_OK)
    return set_err(nullptr, PG_ERR_INVALID, "invalid params: %s", msg);
  if (!d_base || !d_graph || !d_out || (n_query && !d_queries))
    return set_err(nullptr, PG_ERR_INVALID, "NULL argument");
  if (p->precision != PG_FP32)
    return set_err(nullptr, PG_ERR_INVALID,
                   "pg_search_noalloc runs the exact FP32 search only (the "
                   "screened searches read a per-build copy of the base)");
  if (p->Kd > graph_deg)
    return set_err(nullptr, PG_ERR_INVALID, "Kd (%u) exceeds graph degree (%u)",
                   p->Kd, graph_deg);
  const bool persistent = (p->tune_flags & kTunePersistent) != 0;
  const bool ordered = persistent && (p->tune_flags & kTuneOrder) != 0;
  if (ordered && n_query > kNoAllocMaxQueries)
    return set_err(nullptr, PG_ERR_INVALID,
                   "pg_search_noalloc: n_query %u exceeds the %u queries its "
                   "scratch holds for the locality order (clear tune_flags "
                   "bit 9 for larger batches)",
                   n_query, kNoAllocMaxQueries);
  if (n_query == 0) return PG_OK;
  int dev = 0;
  {
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
      return set_err(nullptr, PG_ERR_NO_DEVICE, "hipGetDevice: %s",
                     hipGetErrorString(e));
  }
  uint32_t* ws = nullptr;
  PG_TRY(scratch_address(dev, &ws));
  uint32_t* queue = ws;
  uint32_t* cnt = ws + kScratchQueueWords;
  uint32_t* group = cnt + kScratchCntWords;
  uint32_t* order = group + kNoAllocMaxQueries;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (ordered) {
    const hipError_t e = pg::launch_locality_order(d_queries, n_query, p->dim,
                                                   group, cnt, order, s);
    if (e != hipSuccess)
      return set_err(nullptr, PG_ERR_HIP, "locality order: %s",
                     hipGetErrorString(e));
  }
  const hipError_t e = pg::launch_search(
      d_queries, n_query, d_base, nullptr, nullptr, nullptr, 0, p->n_base,
      p->dim, d_graph, graph_deg, p->Km, p->Kp, p->Kd, p->topk, p->reset_iter,
      p->entry_seed, p->max_iters, effective_table_size(*p), tune_gather(*p),
      d_out, nullptr, persistent ? queue : nullptr, ordered ? order : nullptr,
      tune_waves(*p), s);
  if (e != hipSuccess)
    return set_err(nullptr, PG_ERR_HIP, "search launch: %s",
                   hipGetErrorString(e));
  return PG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Exact top-k gt (pg_exact_topk_ip).
extern "C" int pg_exact_topk_ip(const float* d_queries, uint32_t n_query,
                                const float* d_base, uint32_t n_base,
                                uint32_t dim, uint32_t k, int32_t* d_out,
                                uint32_t* n_exact, void* stream) {
  if (n_exact) *n_exact = 0;
  if ((n_query && (!d_queries || !d_out)) || !d_base)
    return set_err(nullptr, PG_ERR_INVALID, "NULL argument");
  if (dim == 0 || dim > 512)
    return set_err(nullptr, PG_ERR_INVALID, "dim must be in [1, 512]");
  if (k == 0 || k > 256)
    return set_err(nullptr, PG_ERR_INVALID, "k must be in [1, 256]");
  if (n_base == 0 || n_base >= 0x80000000u)
    return set_err(nullptr, PG_ERR_INVALID, "n_base must be in [1, 2^31)");
  if (n_query == 0) return PG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t dpad = (dim + 31) / 32 * 32;
  const uint32_t chunk = n_query < (1u << 20) ? n_query : (1u << 20);
  // accumulation (dpad terms, 2^-23) and chain (dim terms, 2^-24) factors
  const double ua = std::ldexp(1.0, -23), uc = std::ldexp(1.0, -24);
  const float gam_a = float(dpad * ua / (1.0 - dpad * ua));
  const float gam_c = float(dim * uc / (1.0 - dim * uc));
  uint16_t* q16 = nullptr;
  uint16_t* b16 = nullptr;
  float* qn = nullptr;
  float* bn = nullptr;
  float* qr = nullptr;
  float* br = nullptr;
  uint2* buf = nullptr;
  uint32_t* cnt = nullptr;
  uint32_t* list = nullptr;
  uint32_t* nlist = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(s);
    for (void* p : {static_cast<void*>(q16), static_cast<void*>(b16),
                    static_cast<void*>(qn), static_cast<void*>(bn),
                    static_cast<void*>(qr), static_cast<void*>(br),
                    static_cast<void*>(buf), static_cast<void*>(cnt),
                    static_cast<void*>(list), static_cast<void*>(nlist)})
      if (p) (void)hipFree(p);
  };
  auto alloc = [&](auto** ptr, uint64_t bytes, const char* what) -> int {
    if (hipMalloc(reinterpret_cast<void**>(ptr), bytes) != hipSuccess) {
      *ptr = nullptr;
      return set_err(nullptr, PG_ERR_OOM, "hipMalloc(%s, %llu bytes)", what,
                     (unsigned long long)bytes);
    }
    return PG_OK;
  };
  int rc = PG_OK;
  if ((rc = alloc(&q16, uint64_t(chunk) * dpad * 2, "queries_bf16")) ||
      (rc = alloc(&b16, uint64_t(n_base) * dpad * 2, "base_bf16")) ||
      (rc = alloc(&qn, uint64_t(chunk) * 4, "query_norms")) ||
      (rc = alloc(&bn, uint64_t(n_base) * 4, "base_norms")) ||
      (rc = alloc(&qr, uint64_t(chunk) * 4, "query_residual_norms")) ||
      (rc = alloc(&br, uint64_t(n_base) * 4, "base_residual_norms")) ||
      (rc = alloc(&buf, uint64_t(chunk) * pg::kTopkBufferPairs * 8,
                  "candidates")) ||
      (rc = alloc(&cnt, uint64_t(chunk) * 4, "counts")) ||
      (rc = alloc(&list, uint64_t(chunk) * 4, "exact_rows")) ||
      (rc = alloc(&nlist, 4, "exact_row_count"))) {
    release();
    return rc;
  }
  hipError_t e = pg::launch_prepare_screen(d_base, n_base, dim, dpad, b16,
                                           nullptr, bn, br, s);
  uint64_t exact_rows = 0;
  for (uint32_t q0 = 0; e == hipSuccess && q0 < n_query; q0 += chunk) {
    const uint32_t nq = n_query - q0 < chunk ? n_query - q0 : chunk;
    e = pg::launch_prepare_screen(d_queries + uint64_t(q0) * dim, nq, dim,
                                  dpad, q16, nullptr, qn, qr, s);
    if (e == hipSuccess)
      e = pg::launch_exact_topk(d_queries + uint64_t(q0) * dim, nq, d_base,
                                n_base, dim, k, q16, qn, qr, b16, bn, br,
                                dpad, gam_a, gam_c, buf, cnt, list, nlist,
                                d_out + uint64_t(q0) * k, s);
    uint32_t nl = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&nl, nlist, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    exact_rows += nl;
  }
  release();
  if (e != hipSuccess)
    return set_err(nullptr, PG_ERR_HIP, "exact top-k: %s", hipGetErrorString(e));
  if (n_exact) *n_exact = uint32_t(exact_rows);
  return PG_OK;
}
