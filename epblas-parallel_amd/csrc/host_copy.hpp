// host_copy.hpp -- parallel strided host memcpy (host-pointer staging pipeline).
#pragma once
#include <cstddef>
#include <cstdint>

namespace mw {

// dst + c*dpitch <- src + c*spitch, seg bytes per column, c in [0, cols), split over the
// process-wide host worker pool (MWGEMM_HOST_THREADS, default min(16, hardware threads)).
void host_copy_columns(void *dst, size_t dpitch, const void *src, size_t spitch, size_t seg,
                       int64_t cols);

}  // namespace mw
