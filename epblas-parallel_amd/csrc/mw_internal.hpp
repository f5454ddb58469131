// mw_internal.hpp -- internal interfaces between the C-ABI host layer (api.cpp) and the
// kernel translation units.  Nothing here is exported.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "dd_device.hpp"

namespace mw {

enum { OP_N = 0, OP_T = 1, OP_C = 2 };

// One real DD GEMM on device memory: C := alpha*op(A)*op(B) + beta*C, column-major.
struct GemmArgsR {
    int64_t m, n, k;
    int opa, opb;            // OP_N / OP_T (real 'C' folds to OP_T)
    ddv alpha, beta;
    int beta_mode;           // BETA_ZERO / BETA_ONE / BETA_SCALE
    const ddv *A; int64_t lda;
    const ddv *B; int64_t ldb;
    ddv *C; int64_t ldc;
};

struct GemmArgsZ {
    int64_t m, n, k;
    int opa, opb;            // OP_N / OP_T / OP_C
    zddv alpha, beta;
    int beta_mode;
    const zddv *A; int64_t lda;
    const zddv *B; int64_t ldb;
    zddv *C; int64_t ldc;
};

// plain_gemm.hip
hipError_t plain_gemm_real(const GemmArgsR &g, hipStream_t s);
hipError_t plain_gemm_cplx(const GemmArgsZ &g, hipStream_t s);
hipError_t plain_beta_only_real(const GemmArgsR &g, hipStream_t s);
hipError_t plain_beta_only_cplx(const GemmArgsZ &g, hipStream_t s);

}  // namespace mw
