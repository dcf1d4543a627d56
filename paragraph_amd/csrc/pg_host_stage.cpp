// Host stages that complete a ParaGraph index around the GPU stage: the
// entry point, the supply graphs, the bipartite graph, the final lists and
// the index file (C ABI in include/paragraph_host.h, which states the
// canonical rules C4, C6-C8 this file implements).
//
// The reference runs these with OpenMP on a CPU thread next to the GPU
// stage (tests/test_t2iend.cu:888-962, :1228-1258).  Here every list lives
// in flat arrays (CSR or fixed-stride rows) instead of vectors of vectors and
// hash sets, the reverse lists are built by counting in query order, and
// per-node work runs in OpenMP loops with one scratch buffer per thread.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../../include/paragraph.h"
#include "../../include/paragraph_host.h"

namespace {

char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int nthreads(const pg_host_params* p) {
  return p->threads ? int(p->threads) : omp_get_max_threads();
}

// The reference CPU distance (tests/test_t2iend.cu:31-78): -dot over 16 FMA
// lanes (blocks of 16), folded to 8 lanes (+ one block of 8), to 4 lanes
// (+ one block of 4), then the masked remainder on lanes 1..r (rule C8),
// then ((l0 + l1) + (l2 + l3)).
inline float cpu_dist(const float* a, const float* b, uint32_t n) {
  float s16[16];
  for (int j = 0; j < 16; ++j) s16[j] = 0.0f;
  uint32_t i = 0;
  for (; n - i >= 16; i += 16)
    for (int j = 0; j < 16; ++j) s16[j] = std::fmaf(a[i + j], b[i + j], s16[j]);
  float s8[8];
  for (int j = 0; j < 8; ++j) s8[j] = s16[8 + j] + s16[j];
  if (n - i >= 8) {
    for (int j = 0; j < 8; ++j) s8[j] = std::fmaf(a[i + j], b[i + j], s8[j]);
    i += 8;
  }
  float s4[4];
  for (int j = 0; j < 4; ++j) s4[j] = s8[4 + j] + s8[j];
  if (n - i >= 4) {
    for (int j = 0; j < 4; ++j) s4[j] = std::fmaf(a[i + j], b[i + j], s4[j]);
    i += 4;
  }
  const uint32_t r = n - i;
  if (r > 0) {
    for (uint32_t j = 0; j < 4; ++j) {
      const bool on = j >= 1 && j <= r && i + j < n;
      s4[j] = std::fmaf(on ? a[i + j] : 0.0f, on ? b[i + j] : 0.0f, s4[j]);
    }
  }
  return -((s4[0] + s4[1]) + (s4[2] + s4[3]));
}

struct Nb {
  float d;
  uint32_t id;
};

// SimpleNeighbor order (tests/test_t2iend.cu:80-97): distance, then id.
inline bool nb_less(const Nb& a, const Nb& b) {
  return a.d < b.d || (a.d == b.d && a.id < b.id);
}

inline bool contains(const uint32_t* v, size_t n, uint32_t x) {
  for (size_t i = 0; i < n; ++i)
    if (v[i] == x) return true;
  return false;
}

// RNGPrune (tests/test_t2iend.cu:99-177) over a sorted candidate list with
// distinct ids.  Its second pass re-tests every entry against a superset of
// the same kept list, in the same order, so it keeps nothing new and is not
// run.  Non-strict lists are then filled from position 1 on.
void rng_prune(const float* base, uint32_t dim, uint32_t n_base,
               const Nb* fs, uint32_t n, uint32_t target, uint32_t M,
               bool strict, std::vector<uint32_t>& out) {
  out.clear();
  uint32_t s = 0;
  while (s < n && fs[s].id == target) ++s;
  if (s == n) return;
  out.push_back(fs[s].id);
  while (out.size() < M && ++s < n) {
    const Nb& c = fs[s];
    const float* xc = base + size_t(c.id) * dim;
    bool occluded = false;
    for (uint32_t k : out) {
      if (c.id == k || cpu_dist(xc, base + size_t(k) * dim, dim) < c.d) {
        occluded = true;
        break;
      }
    }
    if (!occluded && c.id != target) out.push_back(c.id);
  }
  if (!strict) {
    for (uint32_t j = 1; j < n && out.size() < M; ++j) {
      const uint32_t id = fs[j].id;
      if (id < n_base && id != target && !contains(out.data(), out.size(), id))
        out.push_back(id);
    }
  }
}

int check(const pg_host_params* p) {
  if (!p) return fail(PG_ERR_INVALID, "params is NULL");
  if (p->n_base == 0 || p->dim == 0 || p->gt_k == 0)
    return fail(PG_ERR_INVALID, "n_base, dim and gt_k must be > 0");
  if (p->M_nn == 0 || p->M_nn >= 128)
    return fail(PG_ERR_INVALID, "M_nn must be in [1, 127] (reverse cap 128 - M_nn)");
  if (p->M_final == 0) return fail(PG_ERR_INVALID, "M_final must be > 0");
  if (p->rounds >= p->gt_k)
    return fail(PG_ERR_INVALID, "rounds must be < gt_k (gt columns 1..rounds)");
  return PG_OK;
}

// Writes fixed-stride rows (cnt[i] ids at rows + i * stride) as CSR.
void rows_to_csr(const uint32_t* rows, const uint32_t* cnt, uint32_t n,
                 uint32_t stride, uint32_t* off, uint32_t* ids, int nt) {
  off[0] = 0;
  for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < int64_t(n); ++i)
    std::memcpy(ids + off[i], rows + size_t(i) * stride,
                sizeof(uint32_t) * cnt[i]);
}

// MatchSup round `topn` into fixed-stride rows (stride M_nn).
void match_sup_rows(const pg_host_params* p, const float* base,
                    const uint32_t* gt, uint32_t topn, uint32_t* rows,
                    uint32_t* cnt) {
  const uint32_t N = p->n_base, Q = p->n_query, K = p->gt_k, M = p->M_nn;
  const uint32_t dim = p->dim, cap = 128u - M;
  const int nt = nthreads(p);
  // first-come match on column topn, queries in order (:376-390)
  std::vector<uint32_t> owner(N, UINT32_MAX);  // query that matched base b
  for (uint32_t q = 0; q < Q; ++q) {
    const uint32_t nn = gt[size_t(q) * K + topn];
    if (nn < N && owner[nn] == UINT32_MAX) owner[nn] = q;
  }
  // forward lists: each matched base, its query's gt row, RNG non-strict
  // (:405-428); one row per base (a base is matched at most once)
  std::vector<uint32_t> fwd(size_t(N) * M), fcnt(N, 0);
#pragma omp parallel num_threads(nt)
  {
    std::vector<Nb> fs;
    std::vector<uint32_t> kept;
    fs.reserve(K);
#pragma omp for schedule(dynamic, 256)
    for (int64_t bi = 0; bi < int64_t(N); ++bi) {
      const uint32_t b = uint32_t(bi);
      const uint32_t q = owner[b];
      if (q == UINT32_MAX) continue;
      const uint32_t* g = gt + size_t(q) * K;
      const float* xb = base + size_t(b) * dim;
      fs.clear();
      for (uint32_t j = 0; j < K; ++j) {
        const uint32_t id = g[j];
        if (id >= N) break;
        if (id == b) continue;
        bool dup = false;
        for (const Nb& e : fs)
          if (e.id == id) {
            dup = true;
            break;
          }
        if (dup) continue;
        fs.push_back({cpu_dist(base + size_t(id) * dim, xb, dim), id});
      }
      std::sort(fs.begin(), fs.end(), nb_less);
      rng_prune(base, dim, N, fs.data(), uint32_t(fs.size()), b, M, false,
                kept);
      std::copy(kept.begin(), kept.end(), fwd.begin() + size_t(b) * M);
      fcnt[b] = uint32_t(kept.size());
    }
  }
  // reverse lists b -> sources, in query order, the first `cap` kept (C6)
  std::vector<uint32_t> roff(size_t(N) + 1, 0), rsrc;
  for (uint32_t b = 0; b < N; ++b)
    for (uint32_t t = 0; t < fcnt[b]; ++t) ++roff[fwd[size_t(b) * M + t] + 1];
  for (uint32_t b = 0; b < N; ++b) roff[b + 1] += roff[b];
  rsrc.resize(roff[N]);
  {
    std::vector<uint32_t> fill(roff.begin(), roff.end() - 1);
    for (uint32_t q = 0; q < Q; ++q) {
      const uint32_t b = gt[size_t(q) * K + topn];
      if (b >= N || owner[b] != q) continue;
      for (uint32_t t = 0; t < fcnt[b]; ++t)
        rsrc[fill[fwd[size_t(b) * M + t]]++] = b;
    }
  }
  // merge forward and reverse; over M entries, RNG strict (:450-475)
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> u, kept;
    std::vector<Nb> fs;
#pragma omp for schedule(dynamic, 256)
    for (int64_t bi = 0; bi < int64_t(N); ++bi) {
      const uint32_t b = uint32_t(bi);
      u.assign(fwd.begin() + size_t(b) * M,
               fwd.begin() + size_t(b) * M + fcnt[b]);
      const uint32_t nr = std::min(roff[b + 1] - roff[b], cap);
      u.insert(u.end(), rsrc.begin() + roff[b], rsrc.begin() + roff[b] + nr);
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      uint32_t* row = rows + size_t(b) * M;
      if (u.size() <= M) {  // C7: ascending ids
        std::copy(u.begin(), u.end(), row);
        cnt[b] = uint32_t(u.size());
        continue;
      }
      const float* xb = base + size_t(b) * dim;
      fs.clear();
      for (uint32_t id : u)
        fs.push_back({cpu_dist(base + size_t(id) * dim, xb, dim), id});
      std::sort(fs.begin(), fs.end(), nb_less);
      rng_prune(base, dim, N, fs.data(), uint32_t(fs.size()), b, M, true,
                kept);
      std::copy(kept.begin(), kept.end(), row);
      cnt[b] = uint32_t(kept.size());
    }
  }
}

}  // namespace

extern "C" {

int pg_host_params_init(pg_host_params* p, uint32_t n_base, uint32_t dim,
                        uint32_t n_query) {
  if (!p) return fail(PG_ERR_INVALID, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->n_base = n_base;
  p->dim = dim;
  p->n_query = n_query;
  p->gt_k = 128;
  p->M_nn = 40;
  p->rounds = 6;
  p->M_final = 70;
  return PG_OK;
}

const char* pg_host_last_error(void) { return g_err; }

int pg_host_entry_point(const pg_host_params* p, const uint32_t* gt,
                        uint32_t* ep) {
  int rc = check(p);
  if (rc) return rc;
  if (!ep || (p->n_query && !gt)) return fail(PG_ERR_INVALID, "NULL argument");
  const uint32_t N = p->n_base, K = p->gt_k;
  std::vector<uint32_t> freq(N, 0);
  std::vector<uint8_t> seen(N, 0);
  // each query counts its gt ids up to and including the first one not
  // seen before, which it takes (:189-206)
  for (uint32_t q = 0; q < p->n_query; ++q) {
    const uint32_t* g = gt + size_t(q) * K;
    for (uint32_t j = 0; j < K; ++j) {
      const uint32_t id = g[j];
      if (id >= N) break;
      ++freq[id];
      if (!seen[id]) {
        seen[id] = 1;
        break;
      }
    }
  }
  uint32_t best = 0;
  for (uint32_t b = 1; b < N; ++b)
    if (freq[b] > freq[best]) best = b;  // C4: ties to the smallest id
  *ep = best;
  return PG_OK;
}

int pg_host_match_sup(const pg_host_params* p, const float* base,
                      const uint32_t* gt, uint32_t topn, uint32_t* off,
                      uint32_t* ids) {
  int rc = check(p);
  if (rc) return rc;
  if (!base || !off || !ids || (p->n_query && !gt))
    return fail(PG_ERR_INVALID, "NULL argument");
  if (topn >= p->gt_k) return fail(PG_ERR_INVALID, "topn must be < gt_k");
  std::vector<uint32_t> rows(size_t(p->n_base) * p->M_nn), cnt(p->n_base);
  match_sup_rows(p, base, gt, topn, rows.data(), cnt.data());
  rows_to_csr(rows.data(), cnt.data(), p->n_base, p->M_nn, off, ids,
              nthreads(p));
  return PG_OK;
}

int pg_host_bipartite(const pg_host_params* p, const float* base,
                      const uint32_t* gt, uint32_t* off, uint32_t* ids) {
  int rc = check(p);
  if (rc) return rc;
  if (!base || !off || !ids || (p->n_query && !gt))
    return fail(PG_ERR_INVALID, "NULL argument");
  const uint32_t N = p->n_base, M = p->M_nn, R = p->rounds, dim = p->dim;
  const int nt = nthreads(p);
  // supply graphs of gt columns 1..R (:913-926), fixed-stride rows
  std::vector<uint32_t> rows(size_t(R) * N * M), cnt(size_t(R) * N);
  for (uint32_t r = 0; r < R; ++r)
    match_sup_rows(p, base, gt, r + 1, rows.data() + size_t(r) * N * M,
                   cnt.data() + size_t(r) * N);
  // FusionNN (:490-537): union of the supply lists in round order, without
  // the node itself, RNG strict to 2 * M_nn
  const uint32_t W = 2 * M;
  std::vector<uint32_t> out(size_t(N) * W), ocnt(N);
#pragma omp parallel num_threads(nt)
  {
    std::vector<Nb> fs;
    std::vector<uint32_t> u, kept;
#pragma omp for schedule(dynamic, 256)
    for (int64_t bi = 0; bi < int64_t(N); ++bi) {
      const uint32_t b = uint32_t(bi);
      u.clear();
      for (uint32_t r = 0; r < R; ++r) {
        const uint32_t* row = rows.data() + (size_t(r) * N + b) * M;
        const uint32_t c = cnt[size_t(r) * N + b];
        for (uint32_t t = 0; t < c; ++t)
          if (row[t] < N && row[t] != b) u.push_back(row[t]);
      }
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      const float* xb = base + size_t(b) * dim;
      fs.clear();
      for (uint32_t id : u)
        fs.push_back({cpu_dist(xb, base + size_t(id) * dim, dim), id});
      std::sort(fs.begin(), fs.end(), nb_less);
      rng_prune(base, dim, N, fs.data(), uint32_t(fs.size()), b, W, true,
                kept);
      std::copy(kept.begin(), kept.end(), out.begin() + size_t(b) * W);
      ocnt[b] = uint32_t(kept.size());
    }
  }
  rows_to_csr(out.data(), ocnt.data(), N, W, off, ids, nt);
  return PG_OK;
}

uint32_t pg_host_final_width(const pg_host_params* p) {
  return p ? 2 * p->M_nn + p->M_final : 0;
}

int pg_host_fusion_final(const pg_host_params* p, const float* base,
                         const uint32_t* bi_off, const uint32_t* bi_ids,
                         const uint32_t* link, uint32_t M_link,
                         const uint32_t* proj, uint32_t proj_deg,
                         uint32_t* off, uint32_t* ids) {
  int rc = check(p);
  if (rc) return rc;
  if (!base || !bi_off || !off || !ids || (M_link && !link) ||
      (proj_deg && !proj))
    return fail(PG_ERR_INVALID, "NULL argument");
  const uint32_t N = p->n_base, dim = p->dim, Mf = p->M_final;
  const uint32_t W = pg_host_final_width(p);
  for (uint32_t b = 0; b < N; ++b)
    if (bi_off[b + 1] < bi_off[b] || bi_off[b + 1] - bi_off[b] > 2 * p->M_nn)
      return fail(PG_ERR_INVALID,
                  "bipartite list %u is longer than 2 * M_nn or unordered", b);
  const int nt = nthreads(p);
  std::vector<uint32_t> out(size_t(N) * W), ocnt(N);
#pragma omp parallel num_threads(nt)
  {
    std::vector<Nb> fs;
    std::vector<uint32_t> seen, kept;
#pragma omp for schedule(dynamic, 256)
    for (int64_t bi = 0; bi < int64_t(N); ++bi) {
      const uint32_t b = uint32_t(bi);
      const float* xb = base + size_t(b) * dim;
      uint32_t* row = out.data() + size_t(b) * W;
      uint32_t n_row = 0;
      fs.clear();
      seen.clear();
      auto add = [&](uint32_t id) {
        seen.push_back(id);
        fs.push_back({cpu_dist(xb, base + size_t(id) * dim, dim), id});
      };
      // bipartite entries: kept as they are, and candidates (:553-562)
      for (uint32_t t = bi_off[b]; t < bi_off[b + 1]; ++t) {
        const uint32_t id = bi_ids[t];
        if (id >= N || id == b || contains(seen.data(), seen.size(), id))
          continue;
        add(id);
        row[n_row++] = id;
      }
      // link rows: the node itself is not excluded here (:563-571)
      for (uint32_t j = 0; j < M_link; ++j) {
        const uint32_t id = link[size_t(b) * M_link + j];
        if (id >= N || contains(seen.data(), seen.size(), id)) continue;
        add(id);
      }
      // projection rows (:572-582)
      for (uint32_t j = 0; j < proj_deg; ++j) {
        const uint32_t id = proj[size_t(b) * proj_deg + j];
        if (id >= N || id == b || contains(seen.data(), seen.size(), id))
          continue;
        add(id);
      }
      std::sort(fs.begin(), fs.end(), nb_less);
      rng_prune(base, dim, N, fs.data(), uint32_t(fs.size()), b, Mf, true,
                kept);
      // room left in size_t, as the reference computes it (:596): a
      // bipartite list longer than M_final leaves "room" for every entry
      const size_t room = size_t(Mf) - size_t(n_row);
      const uint32_t n_bi = n_row;
      size_t added = 0;
      for (uint32_t c : kept) {
        if (added >= room) break;
        if (!contains(row, n_bi, c)) {
          row[n_row++] = c;
          ++added;
        }
      }
      ocnt[b] = n_row;
    }
  }
  rows_to_csr(out.data(), ocnt.data(), N, W, off, ids, nt);
  return PG_OK;
}

int pg_host_save_index(const char* path, uint32_t ep, uint32_t n,
                       const uint32_t* off, const uint32_t* ids) {
  if (!path || !off || (n && !ids)) return fail(PG_ERR_INVALID, "NULL argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PG_ERR_INVALID, "cannot open %s for writing", path);
  bool ok = std::fwrite(&ep, 4, 1, f) == 1 && std::fwrite(&n, 4, 1, f) == 1;
  for (uint32_t i = 0; ok && i < n; ++i) {
    const uint32_t c = off[i + 1] - off[i];
    ok = std::fwrite(&c, 4, 1, f) == 1 &&
         (c == 0 || std::fwrite(ids + off[i], 4, c, f) == c);
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PG_OK : fail(PG_ERR_INVALID, "write to %s failed", path);
}

int pg_host_bin_shape(const char* path, uint32_t* rows, uint32_t* cols) {
  if (!path || !rows || !cols) return fail(PG_ERR_INVALID, "NULL argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PG_ERR_INVALID, "cannot open %s", path);
  uint32_t hdr[2];
  const bool ok = std::fread(hdr, 4, 2, f) == 2;
  std::fclose(f);
  if (!ok) return fail(PG_ERR_INVALID, "%s: no 8-byte header", path);
  *rows = hdr[0];
  *cols = hdr[1];
  return PG_OK;
}

int pg_host_read_bin(const char* path, int headered, uint32_t cols,
                     uint64_t row0, uint64_t nrows, void* out) {
  if (!path || (nrows && !out)) return fail(PG_ERR_INVALID, "NULL argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PG_ERR_INVALID, "cannot open %s", path);
  uint64_t file_rows = 0;
  long long base_off = 0;
  if (headered) {
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2) {
      std::fclose(f);
      return fail(PG_ERR_INVALID, "%s: no 8-byte header", path);
    }
    if (cols != 0 && cols != hdr[1]) {
      std::fclose(f);
      return fail(PG_ERR_INVALID, "%s has %u columns, not %u", path, hdr[1],
                  cols);
    }
    cols = hdr[1];
    file_rows = hdr[0];
    base_off = 8;
  } else {
    if (cols == 0) {
      std::fclose(f);
      return fail(PG_ERR_INVALID, "a headerless file needs cols");
    }
    if (std::fseek(f, 0, SEEK_END) != 0) {
      std::fclose(f);
      return fail(PG_ERR_INVALID, "cannot seek in %s", path);
    }
    file_rows = uint64_t(std::ftell(f)) / (4ull * cols);
  }
  if (row0 + nrows > file_rows) {
    std::fclose(f);
    return fail(PG_ERR_INVALID, "%s: rows [%llu, %llu) beyond its %llu rows",
                path, (unsigned long long)row0,
                (unsigned long long)(row0 + nrows),
                (unsigned long long)file_rows);
  }
  const uint64_t elems = nrows * cols;
  bool ok = std::fseek(f, long(base_off + row0 * 4ull * cols), SEEK_SET) == 0;
  // in pieces of 64 MiB (fread counts are size_t)
  uint8_t* dst = static_cast<uint8_t*>(out);
  for (uint64_t done = 0; ok && done < elems;) {
    const uint64_t take = std::min<uint64_t>(elems - done, 16ull << 20);
    ok = std::fread(dst + 4 * done, 4, size_t(take), f) == take;
    done += take;
  }
  std::fclose(f);
  return ok ? PG_OK : fail(PG_ERR_INVALID, "short read from %s", path);
}

int pg_host_pad_gt(const int32_t* in, uint64_t rows, uint32_t in_cols,
                   uint32_t out_cols, int32_t* out) {
  if ((rows && (!in || !out)) || out_cols == 0)
    return fail(PG_ERR_INVALID, "NULL argument or zero width");
  const uint32_t k = in_cols < out_cols ? in_cols : out_cols;
  for (uint64_t r = 0; r < rows; ++r) {
    std::memcpy(out + r * out_cols, in + r * in_cols, 4ull * k);
    for (uint32_t c = k; c < out_cols; ++c) out[r * out_cols + c] = -1;
  }
  return PG_OK;
}

}  // extern "C"
