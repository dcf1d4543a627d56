// pg_build_index: the reference's whole index build (tests/test_t2iend.cu:
// 850-1258: read base and gt, GPU stage beside the CPU bipartite stage,
// FusionFinal, SaveGraph) as a command-line program on the host C++ front
// end (include/paragraph/Builder.hpp) over the two C ABI libraries.
//
//   pg_build_index BASE.fbin GT.ibin M L INDEX_OUT [LINK_OUT PROJ_OUT]
//                  [--precision fp32|bf16|fp8] [--device D]
//
// BASE.fbin / GT.ibin: Big-ANN headered files (gt rows shorter than 128 are
// padded with -1, as transform_gt_to_128.ipynb does).  INDEX_OUT: the
// SaveGraph layout.  LINK_OUT / PROJ_OUT: the GPU stage's link (N x M) and
// projection (N x 16) rows as headered .ibin files.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paragraph/Builder.hpp"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: pg_build_index BASE.fbin GT.ibin M L INDEX_OUT "
               "[LINK_OUT PROJ_OUT] [--precision fp32|bf16|fp8] "
               "[--device D]\n");
  return 2;
}

bool write_ibin(const std::string& path, const std::vector<uint32_t>& rows,
                uint32_t n, uint32_t cols) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const uint32_t hdr[2] = {n, cols};
  bool ok = std::fwrite(hdr, 4, 2, f) == 2 &&
            std::fwrite(rows.data(), 4, rows.size(), f) == rows.size();
  ok = std::fclose(f) == 0 && ok;
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  uint32_t precision = PG_FP32;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--precision" && i + 1 < argc) {
      const std::string v = argv[++i];
      if (v == "fp32") precision = PG_FP32;
      else if (v == "bf16") precision = PG_BF16;
      else if (v == "fp8") precision = PG_FP8E4M3;
      else return usage();
    } else if (a == "--device" && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else {
      pos.push_back(a);
    }
  }
  if (pos.size() != 5 && pos.size() != 7) return usage();
  const uint32_t M = uint32_t(std::strtoul(pos[2].c_str(), nullptr, 10));
  const uint32_t L = uint32_t(std::strtoul(pos[3].c_str(), nullptr, 10));
  uint32_t n = 0, dim = 0, q = 0, gk = 0;
  if (pg_host_bin_shape(pos[0].c_str(), &n, &dim) != PG_OK ||
      pg_host_bin_shape(pos[1].c_str(), &q, &gk) != PG_OK) {
    std::fprintf(stderr, "pg_build_index: %s\n", pg_host_last_error());
    return 1;
  }
  std::vector<float> base(size_t(n) * dim);
  std::vector<int32_t> gt_raw(size_t(q) * gk);
  if (pg_host_read_bin(pos[0].c_str(), 1, dim, 0, n, base.data()) != PG_OK ||
      pg_host_read_bin(pos[1].c_str(), 1, gk, 0, q, gt_raw.data()) != PG_OK) {
    std::fprintf(stderr, "pg_build_index: %s\n", pg_host_last_error());
    return 1;
  }
  const uint32_t gt_k = 128;
  std::vector<int32_t> gt_pad(size_t(q) * gt_k);
  pg_host_pad_gt(gt_raw.data(), q, gk, gt_k, gt_pad.data());
  std::vector<uint32_t> gt(gt_pad.begin(), gt_pad.end());
  try {
    Gbuilder::Builder<float> b(base, gt, dim, gt_k);
    pg_params p;
    pg_params_init(&p, n, dim, q, M, L);
    p.precision = precision;
    if (pos.size() == 7) {
      const Gbuilder::BuildResult r = b.LaunchBuild(M, L, device, &p);
      if (!write_ibin(pos[5], r.link, n, M) ||
          !write_ibin(pos[6], r.projection, n, p.proj_deg)) {
        std::fprintf(stderr, "pg_build_index: cannot write the rows\n");
        return 1;
      }
    }
    const Gbuilder::Index idx = b.BuildIndex(M, L, device, &p);
    Gbuilder::Builder<float>::SaveIndex(idx, pos[4]);
    std::printf("pg_build_index: n=%u dim=%u M=%u L=%u ep=%u ids=%zu\n", n,
                dim, M, L, idx.ep, idx.ids.size());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "pg_build_index: %s\n", e.what());
    return 1;
  }
  return 0;
}
