// Exact FP32 search kernels for query slices of 32 registers per lane
// (dim <= 1024); the launch templates are in pg_search.hpp.
#include "pg_search.hpp"

namespace pg {

hipError_t launch_search_f32_32(const float* base, uint32_t dim, uint32_t u,
                               PG_SEARCH_PARAMS) {
  (void)u;  // one gather batch at this width
  return launch_f32<32, 4>(base, dim, PG_SEARCH_ARGS);
}

}  // namespace pg
