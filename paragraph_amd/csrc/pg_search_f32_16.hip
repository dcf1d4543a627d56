// Exact FP32 search kernels for query slices of 16 registers per lane
// (dim <= 512); the launch templates are in pg_search.hpp.
#include "pg_search.hpp"

namespace pg {

hipError_t launch_search_f32_16(const float* base, uint32_t dim, uint32_t u,
                               PG_SEARCH_PARAMS) {
  if (u <= 4) return launch_f32<16, 4>(base, dim, PG_SEARCH_ARGS);
  if (u <= 8) return launch_f32<16, 8>(base, dim, PG_SEARCH_ARGS);
  return launch_f32<16, 16>(base, dim, PG_SEARCH_ARGS);
}

}  // namespace pg
