#include "ozaki.hpp"
namespace mw {
bool ozaki_supported(int64_t) { return false; }
int ozaki_gemm_real(const GemmArgsR &, hipStream_t) { return record_error(-4, "ozaki path not built"); }
int ozaki_gemm_cplx(const GemmArgsZ &, hipStream_t) { return record_error(-4, "ozaki path not built"); }
}
