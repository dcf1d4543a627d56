// FP8-e4m3-screened search kernels (8 or 16 chunks of 32 padded elements,
// one or two 16-candidate groups in flight); the launch templates are in
// pg_search.hpp.
#include "pg_search.hpp"

namespace pg {

hipError_t launch_search_fp8(const void* base8, const float* scale8,
                             uint32_t dim, uint32_t dpad, bool two,
                             PG_SEARCH_PARAMS) {
  const uint32_t nch = dpad / 32;
  if (nch <= 8)
    return two ? launch_fp8<8, 2>(base8, scale8, dim, dpad, PG_SEARCH_ARGS)
               : launch_fp8<8, 1>(base8, scale8, dim, dpad, PG_SEARCH_ARGS);
  if (nch <= 16)
    return two ? launch_fp8<16, 2>(base8, scale8, dim, dpad, PG_SEARCH_ARGS)
               : launch_fp8<16, 1>(base8, scale8, dim, dpad, PG_SEARCH_ARGS);
  return hipErrorInvalidValue;
}

}  // namespace pg
