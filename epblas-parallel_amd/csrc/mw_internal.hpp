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
    int variant = -1;        // tile variant; -1: autotuned choice or heuristic
    int order = 0;           // plain path: ORDER_SIMD or ORDER_SCALAR (plain_tile.hpp)
    // A call of the drop-in entry points (mgemm_, mgemm_64_, mgemm_serial), whose operands are
    // staged in one piece or in column panels: its exact product keeps the worst-case
    // modulus count, because column panels would each choose a count from their own
    // columns, and a different count rounds the reconstruction differently; so a call gives
    // the same bits however it is staged.
    bool host_call = false;
};

struct GemmArgsZ {
    int64_t m, n, k;
    int opa, opb;            // OP_N / OP_T / OP_C
    zddv alpha, beta;
    int beta_mode;
    const zddv *A; int64_t lda;
    const zddv *B; int64_t ldb;
    zddv *C; int64_t ldc;
    int variant = -1;
    int order = 0;
    // Triangular-output GEMM only: bit 0 / bit 1 = the diagonal triangles take op(A) /
    // op(B) as OP_T where the rectangle has OP_C (a letter outside N/T/C, which the
    // reference's diagonal step and GEMM packers read differently).
    int diag_alt = 0;
};

// Operation orders of the plain path: the reference's AVX2 kernel, or its scalar kernel
// (hosts without AVX2/FMA).  Chosen per call from the library setting (api.cpp).
enum PlainOrder { ORDER_SIMD = 0, ORDER_SCALAR = 1 };
int plain_order();   // the current setting

// Device buffers the library allocates for itself (exact-product workspace, host staging,
// batched plans) go through these (api.cpp).  mwgemm_testing_set_alloc_limit makes every
// such buffer above a size unobtainable, as on a device without that much free memory.
hipError_t device_alloc(void **p, size_t bytes);
bool device_alloc_allowed(size_t bytes);

// Device-pointer GEMM cores (api.cpp).  a_tag: see ozaki_gemm_real.
int gemm_real_device(GemmArgsR g, int path, hipStream_t s, uint64_t a_tag = 0,
                     bool noalloc = false);
int gemm_cplx_device(GemmArgsZ g, int path, hipStream_t s, uint64_t a_tag = 0);

// Autotuner table (autotune.cpp).  Kinds of tuned decisions:
enum TuneKind {
    TK_PLAIN_REAL = 0,   // variant: 0 = 64x64 tile, 1 = 32x64 tile
    TK_PLAIN_CPLX = 1,   // variant: 0 = 32x32 tile
    TK_OZ_GEMM = 2,      // variant: residue tile 0 = 256x256, 1 = 128x128, 2 = 256x128, 3 = 128x256
    TK_BATCH = 3,        // variant: path * 16 + residue-tile variant (0 = 128, 1 = 64,
                         // 2 = 64 with the reconstruction fused)
    TK_COUNT
};
// Tuned variant for a shape class, or -1 when none has been recorded.
int tune_lookup(int kind, int64_t a, int64_t b, int64_t c);
void tune_record(int kind, int64_t a, int64_t b, int64_t c, int variant, double ms);
constexpr int PLAIN_REAL_VARIANTS = 2;
constexpr int OZ_GEMM_VARIANTS = 4;
constexpr int BATCH_OZ_VARIANTS = 3;

// plain_gemm.hip
hipError_t plain_gemm_real(const GemmArgsR &g, hipStream_t s);
hipError_t plain_gemm_cplx(const GemmArgsZ &g, hipStream_t s);
hipError_t plain_beta_only_real(const GemmArgsR &g, hipStream_t s);
hipError_t plain_beta_only_cplx(const GemmArgsZ &g, hipStream_t s);
// Matrix-vector products (gemv.hip), op in OP_N/OP_T/OP_C (real 'C' = 'T'); nbands: the
// reference's OpenMP row split for real op 'N'; order: ORDER_SIMD or ORDER_SCALAR.
hipError_t gemv_real(int op, int64_t m, int64_t n, ddv alpha, ddv beta, const ddv *A, int64_t lda,
                     const ddv *x, int64_t incx, ddv *y, int64_t incy, int64_t nbands,
                     int order, hipStream_t s);
hipError_t gemv_cplx(int op, int64_t m, int64_t n, zddv alpha, zddv beta, const zddv *A,
                     int64_t lda, const zddv *x, int64_t incx, zddv *y, int64_t incy, int order,
                     hipStream_t s);
// Triangular-output GEMM on the UPLO triangle of the n x n C (g.m == g.n), column blocks of nb.
hipError_t plain_gemmtr_real(const GemmArgsR &g, bool upper, int64_t nb, hipStream_t s);
hipError_t plain_gemmtr_cplx(const GemmArgsZ &g, bool upper, int64_t nb, hipStream_t s);

}  // namespace mw
