// Gbuilder::Builder<T> — host C++ front end of the index-build stage.
//
// Mirrors the reference's (commented-out) class, include/Builder.hpp:6-41:
// constructed from the base vectors and the gt ids, SetGt() replaces the gt,
// copy construction / assignment are deleted.  LaunchBuild() takes the build
// knobs at run time (the reference's were template arguments,
// tests/test_t2iend.cu:670-727) and runs the GPU stage through the C ABI in
// paragraph.h, returning the projection (N x 16) and link (N x M) rows that
// the reference copies back for FusionFinal (tests/test_t2iend.cu:1033,1153).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <thread>

#include "../paragraph.h"
#include "../paragraph_host.h"

namespace Gbuilder {

// The final index: entry point and one list per base as CSR.
struct Index {
  uint32_t ep = 0;
  std::vector<uint32_t> off;  // N + 1
  std::vector<uint32_t> ids;
};

struct BuildResult {
  std::vector<uint32_t> projection;  // N x proj_deg, tomb-padded
  std::vector<uint32_t> link;        // N x M, tomb-padded
  pg_stats stats;
};

template <typename T>
class Builder {
  static_assert(sizeof(T) == 4, "base vectors are float32");

 public:
  using data_type = T;
  using id_type = uint32_t;

  Builder() = delete;
  explicit Builder(std::vector<data_type> const& base,
                   std::vector<id_type> const& gt_ids, uint32_t dim,
                   uint32_t gt_k = 128)
      : base_(base), gt_ids_(gt_ids), dim_(dim), gt_k_(gt_k) {}
  Builder(Builder&& other) noexcept
      : base_(std::move(other.base_)),
        gt_ids_(std::move(other.gt_ids_)),
        dim_(other.dim_),
        gt_k_(other.gt_k_) {}
  Builder(Builder& other) = delete;
  Builder& operator=(Builder const&) = delete;
  Builder& operator=(Builder&&) = delete;
  ~Builder() = default;

  auto SetGt(std::vector<id_type>&& gt_ids) -> void { gt_ids_ = std::move(gt_ids); }

  // Runs the whole GPU stage on `device`.  `params` defaults to the
  // project mapping of (M, L); throws std::runtime_error with the library's
  // message on failure.
  auto LaunchBuild(uint32_t M, uint32_t L, int device = 0,
                   pg_params const* params = nullptr) -> BuildResult {
    if (dim_ == 0 || base_.size() % dim_ != 0)
      throw std::runtime_error("base size is not a multiple of dim");
    const uint32_t n = uint32_t(base_.size() / dim_);
    const uint32_t q = uint32_t(gt_ids_.size() / gt_k_);
    pg_params p;
    if (params) {
      p = *params;
    } else {
      pg_params_init(&p, n, dim_, q, M, L);
      p.gt_k = gt_k_;
    }
    pg_ctx* ctx = pg_create(device, &p);
    if (!ctx) throw std::runtime_error(pg_last_error(nullptr));
    BuildResult r;
    r.projection.resize(size_t(n) * p.proj_deg);
    r.link.resize(size_t(n) * p.M);
    int rc = pg_gpu_stage_async(ctx, reinterpret_cast<const float*>(base_.data()),
                                gt_ids_.data(), r.projection.data(), r.link.data());
    if (rc == PG_OK) rc = pg_wait(ctx);
    if (rc == PG_OK) rc = pg_get_stats(ctx, &r.stats);
    if (rc != PG_OK) {
      std::string msg = pg_last_error(ctx);
      pg_destroy(ctx);
      throw std::runtime_error(msg);
    }
    pg_destroy(ctx);
    return r;
  }

  // The reference's whole build (tests/test_t2iend.cu:850-1258): the GPU
  // stage, with the host bipartite stage (MatchSup rounds + FusionNN) on a
  // CPU thread beside it, then FusionFinal and the entry point.  `hp`
  // defaults to the reference's host values (pg_host_params_init).  Links
  // libparagraph_hip and libparagraph_host.
  auto BuildIndex(uint32_t M, uint32_t L, int device = 0,
                  pg_params const* params = nullptr,
                  pg_host_params const* hp = nullptr) -> Index {
    const uint32_t n = uint32_t(base_.size() / dim_);
    const uint32_t q = uint32_t(gt_ids_.size() / gt_k_);
    pg_host_params h;
    if (hp) {
      h = *hp;
    } else {
      pg_host_params_init(&h, n, dim_, q);
      h.gt_k = gt_k_;
    }
    const float* base = reinterpret_cast<const float*>(base_.data());
    std::vector<uint32_t> bi_off(size_t(n) + 1), bi_ids(size_t(n) * 2 * h.M_nn);
    int host_rc = PG_OK;
    std::thread host([&] {
      host_rc = pg_host_bipartite(&h, base, gt_ids_.data(), bi_off.data(),
                                  bi_ids.data());
    });
    BuildResult r;
    try {
      r = LaunchBuild(M, L, device, params);
    } catch (...) {
      host.join();
      throw;
    }
    host.join();
    if (host_rc != PG_OK) throw std::runtime_error(pg_host_last_error());
    Index idx;
    idx.off.resize(size_t(n) + 1);
    idx.ids.resize(size_t(n) * pg_host_final_width(&h));
    const uint32_t proj_deg = uint32_t(r.projection.size() / n);
    const uint32_t m_link = uint32_t(r.link.size() / n);
    if (pg_host_entry_point(&h, gt_ids_.data(), &idx.ep) != PG_OK ||
        pg_host_fusion_final(&h, base, bi_off.data(), bi_ids.data(),
                             r.link.data(), m_link, r.projection.data(),
                             proj_deg, idx.off.data(), idx.ids.data()) != PG_OK)
      throw std::runtime_error(pg_host_last_error());
    idx.ids.resize(idx.off[n]);
    return idx;
  }

  // SaveGraph layout (tests/test_t2iend.cu:615-633).
  static auto SaveIndex(Index const& idx, std::string const& path) -> void {
    if (pg_host_save_index(path.c_str(), idx.ep, uint32_t(idx.off.size() - 1),
                           idx.off.data(), idx.ids.data()) != PG_OK)
      throw std::runtime_error(pg_host_last_error());
  }

 private:
  std::vector<data_type> base_;
  std::vector<id_type> gt_ids_;
  uint32_t dim_;
  uint32_t gt_k_;
};

}  // namespace Gbuilder
