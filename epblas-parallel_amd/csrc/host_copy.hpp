// host_copy.hpp -- host-only helpers: parallel strided host memcpy (host-pointer staging
// pipeline) and the host CPU feature test.
#pragma once
#include <cstddef>
#include <cstdint>

namespace mw {

// dst + c*dpitch <- src + c*spitch, seg bytes per column, c in [0, cols), split over the
// process-wide host worker pool (MWGEMM_HOST_THREADS, default min(16, hardware threads)).
void host_copy_columns(void *dst, size_t dpitch, const void *src, size_t spitch, size_t seg,
                       int64_t cols);

// Whether the reference would run its AVX2/FMA kernels on this host: it builds them only for
// x86-64 and enters them only when the CPU has both features (mf_cpu.cpp:10-23,
// mf_dispatch.h:17-41); elsewhere its scalar kernels run.
bool host_has_avx2_fma();

}  // namespace mw
