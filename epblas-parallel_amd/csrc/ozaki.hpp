// ozaki.hpp -- the Ozaki-scheme-II double-double GEMM (INT8 residues on the matrix cores).
#pragma once
#include "mw_internal.hpp"

namespace mw {

int record_error(int code, const char *fmt, ...);

// True when the exact-product path handles a contraction depth of k.
bool ozaki_supported(int64_t k);
// C := alpha*op(A)*op(B) + beta*C through the residue pipeline; returns 0 or an error code.
int ozaki_gemm_real(const GemmArgsR &g, hipStream_t s);
int ozaki_gemm_cplx(const GemmArgsZ &g, hipStream_t s);

}  // namespace mw
