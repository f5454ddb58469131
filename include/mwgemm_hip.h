/*
 * mwgemm_hip.h -- C ABI of the MI355X (gfx950) double-double GEMM library
 * (libmwgemm_hip.so).
 *
 * The first group of entry points is the drop-in boundary: the same symbols, argument
 * meaning and calling convention as the reference's mwblas GEMMs, so an existing caller
 * links this library in front of (or instead of) epblas-parallel::mwblas and its mgemm_ /
 * wgemm_ calls run on the GPU.  Matrices are column-major; every argument is passed by
 * reference (Fortran ABI); no hidden character-length arguments are declared.  Pointers may
 * be host memory (staged through the device via double-buffered pinned chunks, copied by
 * a host thread pool) or device memory (used in place).
 *
 * The second group has no counterpart in the reference: device-pointer entry points with
 * an explicit HIP stream, a batched entry point, the numeric-path switch, error reporting
 * and kernel timers.
 */
#ifndef MWGEMM_HIP_H
#define MWGEMM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MWGEMM_API
#define MWGEMM_API __attribute__((visibility("default")))
#endif

/* POD layouts of multifloats::float64x2 {double limbs[2]} and complex64x2 {re, im}
 * (third-party multifloats v0.8.1; mirrored at
 * benchmark/drivers/target_multifloats/perf_mgemm.cpp:13-14). */
typedef struct { double limbs[2]; } mw_float64x2;
typedef struct { mw_float64x2 re, im; } mw_complex64x2;

/* ---------------------------------------------------------------------------------
 * Drop-in boundary.
 *
 * Letters are read as the reference reads them; none is rejected.  Real routines: any
 * letter but 'N' (either case) transposes (common/blas_char.h:33-36, mgemm_serial.cpp:52).
 * Complex op(A): 'N', 'T', anything else conjugate-transposes (wgemm_serial.cpp:56-73).
 * Complex op(B): 'N', 'T', 'C'; another letter transposes in the AVX2 operation order and
 * conjugate-transposes in the scalar order, as the reference's two packers do
 * (wgemm_serial.cpp:145, 174-185 and :81-99).  wgemv: 'N', 'C', anything else transposes.
 * A call whose exact-product workspace cannot be allocated computes on the plain path.
 * ------------------------------------------------------------------------------- */

/* C := alpha*op(A)*op(B) + beta*C over double-double, LP64 integers.
 * Replaces mgemm_ = EPBLAS_FACADE_GEMM(mgemm, TR) at
 * src/epblas-parallel/multifloats/mgemm_parallel.cpp:189-191
 * (macro src/epblas-parallel/common/epblas_facade.h:234-240, LP64 twin :63-65). */
MWGEMM_API void mgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const mw_float64x2 *alpha, const mw_float64x2 *a, const int *lda,
            const mw_float64x2 *b, const int *ldb, const mw_float64x2 *beta,
            mw_float64x2 *c, const int *ldc);

/* ILP64 twin of mgemm_ (epblas_facade.h:63-65, suffix _64_). */
MWGEMM_API void mgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n,
               const int64_t *k, const mw_float64x2 *alpha, const mw_float64x2 *a,
               const int64_t *lda, const mw_float64x2 *b, const int64_t *ldb,
               const mw_float64x2 *beta, mw_float64x2 *c, const int64_t *ldc);

/* Complex double-double GEMM, op letters N/T/C.  Replaces wgemm_ at
 * src/epblas-parallel/multifloats/wgemm_parallel.cpp:196-198 (same facade macro). */
MWGEMM_API void wgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const mw_complex64x2 *alpha, const mw_complex64x2 *a, const int *lda,
            const mw_complex64x2 *b, const int *ldb, const mw_complex64x2 *beta,
            mw_complex64x2 *c, const int *ldc);

/* ILP64 twin of wgemm_. */
MWGEMM_API void wgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n,
               const int64_t *k, const mw_complex64x2 *alpha, const mw_complex64x2 *a,
               const int64_t *lda, const mw_complex64x2 *b, const int64_t *ldb,
               const mw_complex64x2 *beta, mw_complex64x2 *c, const int64_t *ldc);

/* By-value core called by the reference's other mw level-3 routines for their trailing
 * updates (msyrk_serial.cpp:336-352, msymm_serial.cpp:396-455, mtrsm_serial.cpp:772-813).
 * Replaces mgemm_serial, src/epblas-parallel/multifloats/mgemm_kernel.h:68-75. */
MWGEMM_API void mgemm_serial(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                  const mw_float64x2 *alpha, const mw_float64x2 *a, ptrdiff_t lda,
                  const mw_float64x2 *b, ptrdiff_t ldb, const mw_float64x2 *beta,
                  mw_float64x2 *c, ptrdiff_t ldc);

/* Replaces wgemm_serial, src/epblas-parallel/multifloats/wgemm_kernel.h:62-69. */
MWGEMM_API void wgemm_serial(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                  const mw_complex64x2 *alpha, const mw_complex64x2 *a, ptrdiff_t lda,
                  const mw_complex64x2 *b, ptrdiff_t ldb, const mw_complex64x2 *beta,
                  mw_complex64x2 *c, ptrdiff_t ldc);

/* Triangular-output GEMM: C := alpha*op(A)*op(B) + beta*C on the UPLO triangle ('U' or
 * 'L') of the n x n C only; op(A) is n x k, op(B) is k x n.  Replaces mgemmtr_ =
 * EPBLAS_FACADE_GEMMTR(mgemmtr, TR), src/epblas-parallel/multifloats/mgemmtr_parallel.cpp:82
 * (macro common/epblas_facade.h:242-248), whose numerics live in mgemmtr_serial.cpp:117-200.
 * Bit-for-bit the reference's result: columns are walked in blocks of nb (MGEMMTR_NB,
 * default 8), the nb x nb diagonal triangles take the scalar faithful sequence, the rest
 * the GEMM's per-element sequence.  As in the reference, uplo 'U' (either case) selects the
 * upper triangle and any other letter the lower (mgemmtr_serial.cpp:178). */
MWGEMM_API void mgemmtr_(const char *uplo, const char *transa, const char *transb, const int *n,
              const int *k, const mw_float64x2 *alpha, const mw_float64x2 *a, const int *lda,
              const mw_float64x2 *b, const int *ldb, const mw_float64x2 *beta, mw_float64x2 *c,
              const int *ldc);

/* ILP64 twin of mgemmtr_ (epblas_facade.h:63-65). */
MWGEMM_API void mgemmtr_64_(const char *uplo, const char *transa, const char *transb,
                 const int64_t *n, const int64_t *k, const mw_float64x2 *alpha,
                 const mw_float64x2 *a, const int64_t *lda, const mw_float64x2 *b,
                 const int64_t *ldb, const mw_float64x2 *beta, mw_float64x2 *c,
                 const int64_t *ldc);

/* Complex twin (op letters N/T/C, WGEMMTR_NB).  Replaces wgemmtr_ =
 * EPBLAS_FACADE_GEMMTR(wgemmtr, TC), wgemmtr_parallel.cpp:85; numerics wgemmtr_serial.cpp. */
MWGEMM_API void wgemmtr_(const char *uplo, const char *transa, const char *transb, const int *n,
              const int *k, const mw_complex64x2 *alpha, const mw_complex64x2 *a,
              const int *lda, const mw_complex64x2 *b, const int *ldb,
              const mw_complex64x2 *beta, mw_complex64x2 *c, const int *ldc);

/* ILP64 twin of wgemmtr_. */
MWGEMM_API void wgemmtr_64_(const char *uplo, const char *transa, const char *transb,
                 const int64_t *n, const int64_t *k, const mw_complex64x2 *alpha,
                 const mw_complex64x2 *a, const int64_t *lda, const mw_complex64x2 *b,
                 const int64_t *ldb, const mw_complex64x2 *beta, mw_complex64x2 *c,
                 const int64_t *ldc);

/* By-value cores.  Replace mgemmtr_serial (mgemmtr_kernel.h:56-63) and wgemmtr_serial
 * (wgemmtr_kernel.h:52-59). */
MWGEMM_API void mgemmtr_serial(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                    const mw_float64x2 *alpha, const mw_float64x2 *a, ptrdiff_t lda,
                    const mw_float64x2 *b, ptrdiff_t ldb, const mw_float64x2 *beta,
                    mw_float64x2 *c, ptrdiff_t ldc);
MWGEMM_API void wgemmtr_serial(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                    const mw_complex64x2 *alpha, const mw_complex64x2 *a, ptrdiff_t lda,
                    const mw_complex64x2 *b, ptrdiff_t ldb, const mw_complex64x2 *beta,
                    mw_complex64x2 *c, ptrdiff_t ldc);

/* Matrix-vector product: y := alpha*op(A)*x + beta*y, A m x n, op 'N', 'T' ('C' = 'T'),
 * strided x and y (negative strides from the far end).  Replaces mgemv_ =
 * EPBLAS_FACADE_GEMV(mgemv, TR), src/epblas-parallel/multifloats/mgemv.cpp:277-279
 * (macro common/epblas_facade.h:132-138; numerics mgemv.cpp:60-275).  Bit-for-bit the
 * reference's result; for op 'N' that includes its OpenMP row bands (the rows after the last
 * whole 4-block of each band use faithful operations), reproduced for the team size
 * MWGEMV_THREADS, else the OpenMP default team (the first entry of OMP_NUM_THREADS, else the
 * CPUs of the process affinity mask, capped by OMP_THREAD_LIMIT; m < 64: one band).
 * m == 0 or n == 0 leaves y untouched.  Zero increments and short lda are rejected. */
MWGEMM_API void mgemv_(const char *trans, const int *m, const int *n, const mw_float64x2 *alpha,
            const mw_float64x2 *a, const int *lda, const mw_float64x2 *x, const int *incx,
            const mw_float64x2 *beta, mw_float64x2 *y, const int *incy);
/* ILP64 twin of mgemv_. */
MWGEMM_API void mgemv_64_(const char *trans, const int64_t *m, const int64_t *n,
               const mw_float64x2 *alpha, const mw_float64x2 *a, const int64_t *lda,
               const mw_float64x2 *x, const int64_t *incx, const mw_float64x2 *beta,
               mw_float64x2 *y, const int64_t *incy);
/* Complex twin (op N/T/C).  Replaces wgemv_ = EPBLAS_FACADE_GEMV(wgemv, TC), wgemv.cpp; its
 * op-'N' row split is 4-aligned, so only the final m % 4 rows take the faithful tail. */
MWGEMM_API void wgemv_(const char *trans, const int *m, const int *n, const mw_complex64x2 *alpha,
            const mw_complex64x2 *a, const int *lda, const mw_complex64x2 *x, const int *incx,
            const mw_complex64x2 *beta, mw_complex64x2 *y, const int *incy);
/* ILP64 twin of wgemv_. */
MWGEMM_API void wgemv_64_(const char *trans, const int64_t *m, const int64_t *n,
               const mw_complex64x2 *alpha, const mw_complex64x2 *a, const int64_t *lda,
               const mw_complex64x2 *x, const int64_t *incx, const mw_complex64x2 *beta,
               mw_complex64x2 *y, const int64_t *incy);

/* ---------------------------------------------------------------------------------
 * Library extensions (no counterpart in the reference).
 * ------------------------------------------------------------------------------- */

/* Numeric path.  PLAIN reproduces the reference's results bit for bit (FP64 TwoProd/TwoSum
 * on the vector ALU); OZAKI computes the product exactly from INT8 residues on the matrix
 * cores (normwise error below 1e-30, not bitwise equal to PLAIN); AUTO picks OZAKI for
 * large products.  The default comes from the environment variable MWGEMM_PATH
 * ("plain" | "ozaki" | "auto"), else PLAIN. */
enum { MWGEMM_PATH_PLAIN = 0, MWGEMM_PATH_OZAKI = 1, MWGEMM_PATH_AUTO = 2 };
MWGEMM_API int mwgemm_set_path(int path);          /* returns the previous path, -1 on a bad value */
MWGEMM_API int mwgemm_get_path(void);

/* Operation order of the plain path (mgemm_/wgemm_ and their _serial cores, the gemmtr
 * rectangle, batched plain plans, mgemv_/wgemv_).  The reference computes with one of two
 * kernels, chosen on its host: the AVX2/FMA kernels (SIMD: fast operations accumulated in
 * 128-deep chunks, mgemm_serial.cpp:173-319) where the CPU has AVX2 and FMA, else scalar
 * kernels (SCALAR: every multiply-add into C with the faithful operations,
 * mgemm_serial.cpp:87-99, wgemm_serial.cpp:102-115, mgemv.cpp:147-153, :213-223).  The two
 * round differently, so the plain path reproduces whichever the reference would run: by
 * default the same host test the reference makes (mf_have_avx2_fma, mf_cpu.cpp:10-23),
 * overridden by MWGEMM_PLAIN_ORDER ("simd" | "scalar") or this setter. */
enum { MWGEMM_ORDER_SIMD = 0, MWGEMM_ORDER_SCALAR = 1 };
MWGEMM_API int mwgemm_set_plain_order(int order);  /* returns the previous order, -1 on a bad value */
MWGEMM_API int mwgemm_get_plain_order(void);

/* Error handling.  Every failure is reported on stderr and recorded; the drop-in entry
 * points then abort the process unless MWGEMM_ON_ERROR=return (or mwgemm_set_abort(0)),
 * in which case C is left unmodified and the error stays readable here.  The device entry
 * points never abort; they return the error code. */
MWGEMM_API int mwgemm_last_error(char *buf, size_t len);   /* last error code (0 = none), message */
MWGEMM_API void mwgemm_clear_error(void);
MWGEMM_API int mwgemm_set_abort(int on);                  /* returns the previous setting */

/* Device-pointer GEMMs on the given stream; NULL is HIP's default (null) stream, as for
 * every HIP launch, so the call is ordered after work queued there.  alpha/beta are host
 * values.  Return 0 or a hipError_t code.  The call is asynchronous w.r.t. the host; path
 * is one of MWGEMM_PATH_*.  The device-pointer gemmtr, gemv and batched entry points below
 * take their stream the same way, and accept only the letters N/T/C and U/L.
 * Exact-product calls share one workspace per device.  Eager calls on different streams are
 * ordered on the device.  A call captured into a HIP graph (the workspace must already be
 * large enough: make the same call once outside capture first) keeps its buffer for the
 * graph, and later eager calls move to a new one, so graph replays and eager calls never
 * share memory; replays of graphs captured on one device must be ordered by the caller. */
MWGEMM_API int mwgemm_mgemm_device(char transa, char transb, int64_t m, int64_t n, int64_t k,
                        const mw_float64x2 *alpha, const mw_float64x2 *dA, int64_t lda,
                        const mw_float64x2 *dB, int64_t ldb, const mw_float64x2 *beta,
                        mw_float64x2 *dC, int64_t ldc, int path, void *stream);
/* No counterpart in the reference: mwgemm_mgemm_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and stream semantics; it calls no hipMalloc, hipFree
 * or other allocating routine, and kernel timers are not recorded for it.  The plain path
 * needs no memory beyond the operands.  The exact-product path keeps its workspace in a
 * static per-device scratch array of 128 MiB (held by each GPU the library runs on until
 * the process ends, besides 138 KiB of modulus tables in its head); a product whose
 * workspace does not fit (about 1100^3 for a square one) is rejected with an error when
 * OZAKI is requested, and computed on the plain path under AUTO.  The first call on a device
 * sets up its scratch (address lookup, table upload) and must not be captured.  Every call
 * on a GPU shares that GPU's scratch: calls that overlap on one GPU race. */
MWGEMM_API int mwgemm_mgemm_device_noalloc(char transa, char transb, int64_t m, int64_t n,
                                           int64_t k, const mw_float64x2 *alpha,
                                           const mw_float64x2 *dA, int64_t lda,
                                           const mw_float64x2 *dB, int64_t ldb,
                                           const mw_float64x2 *beta, mw_float64x2 *dC,
                                           int64_t ldc, int path, void *stream);
MWGEMM_API int mwgemm_wgemm_device(char transa, char transb, int64_t m, int64_t n, int64_t k,
                        const mw_complex64x2 *alpha, const mw_complex64x2 *dA, int64_t lda,
                        const mw_complex64x2 *dB, int64_t ldb, const mw_complex64x2 *beta,
                        mw_complex64x2 *dC, int64_t ldc, int path, void *stream);

/* Device-pointer triangular-output GEMMs (plain path, asynchronous, status code). */
MWGEMM_API int mwgemm_mgemmtr_device(char uplo, char transa, char transb, int64_t n, int64_t k,
                          const mw_float64x2 *alpha, const mw_float64x2 *dA, int64_t lda,
                          const mw_float64x2 *dB, int64_t ldb, const mw_float64x2 *beta,
                          mw_float64x2 *dC, int64_t ldc, void *stream);
MWGEMM_API int mwgemm_wgemmtr_device(char uplo, char transa, char transb, int64_t n, int64_t k,
                          const mw_complex64x2 *alpha, const mw_complex64x2 *dA, int64_t lda,
                          const mw_complex64x2 *dB, int64_t ldb, const mw_complex64x2 *beta,
                          mw_complex64x2 *dC, int64_t ldc, void *stream);

/* Device-pointer matrix-vector products (asynchronous, status code).  nbands: the row bands
 * of the real op-'N' product (<= 0: the drop-in's rule above). */
MWGEMM_API int mwgemm_mgemv_device(char trans, int64_t m, int64_t n, const mw_float64x2 *alpha,
                        const mw_float64x2 *dA, int64_t lda, const mw_float64x2 *dx,
                        int64_t incx, const mw_float64x2 *beta, mw_float64x2 *dy, int64_t incy,
                        int64_t nbands, void *stream);
MWGEMM_API int mwgemm_wgemv_device(char trans, int64_t m, int64_t n, const mw_complex64x2 *alpha,
                        const mw_complex64x2 *dA, int64_t lda, const mw_complex64x2 *dx,
                        int64_t incx, const mw_complex64x2 *beta, mw_complex64x2 *dy,
                        int64_t incy, void *stream);

/* Batched GEMMs: count independent real DD problems C_p := alpha_p op(A_p) op(B_p) +
 * beta_p C_p (one op pair for the whole batch; per-problem sizes, leading dimensions,
 * scalars and device pointers; all host arrays of length count).  A plan validates the
 * batch, allocates its device table and workspace, and captures the kernel sequence into a
 * HIP graph: the plain path is one persistent work-queue kernel over all output tiles
 * (results bit-identical to separate mgemm_ calls), the Ozaki path a fixed sequence of
 * flattened/persistent kernels.  plan_run replays the graph on a stream; the plan is bound
 * to the pointers it was created with.  Runs of one plan share its work counters and
 * workspace: a run on another stream than the previous run waits for it on the device.  mwgemm_mgemm_batched_device builds, runs (without
 * a graph) and frees a plan in one call and synchronises the stream.  Return 0 or an error
 * code.  No counterpart in the reference (its mgemm_ is one problem per call). */
typedef struct mwgemm_batched_plan_s *mwgemm_batched_plan;
MWGEMM_API int mwgemm_batched_plan_create(mwgemm_batched_plan *plan, int64_t count, char transa,
                                          char transb, const int64_t *m, const int64_t *n,
                                          const int64_t *k, const mw_float64x2 *alpha,
                                          const mw_float64x2 *const *dA, const int64_t *lda,
                                          const mw_float64x2 *const *dB, const int64_t *ldb,
                                          const mw_float64x2 *beta, mw_float64x2 *const *dC,
                                          const int64_t *ldc, int path);
MWGEMM_API int mwgemm_batched_plan_run(mwgemm_batched_plan plan, void *stream);
MWGEMM_API int mwgemm_batched_plan_info(mwgemm_batched_plan plan, int *path, int64_t *tiles,
                                        double *flops, int *nmod);
MWGEMM_API int mwgemm_batched_plan_destroy(mwgemm_batched_plan plan);
MWGEMM_API int mwgemm_mgemm_batched_device(int64_t count, char transa, char transb,
                                           const int64_t *m, const int64_t *n, const int64_t *k,
                                           const mw_float64x2 *alpha,
                                           const mw_float64x2 *const *dA, const int64_t *lda,
                                           const mw_float64x2 *const *dB, const int64_t *ldb,
                                           const mw_float64x2 *beta, mw_float64x2 *const *dC,
                                           const int64_t *ldc, int path, void *stream);

/* Autotuner.  mwgemm_autotune times every tile variant of the given path (PLAIN or OZAKI)
 * on synthetic device operands of shape m x n x k (allocated and freed here) and records the
 * fastest for that shape class; mwgemm_batched_autotune does the same for a batch (plain
 * vs Ozaki with 128 x 128 residue tiles (variant 0), 64 x 64 tiles (1), or 64 x 64 tiles
 * with the reconstruction fused into them (2)), and plans created with path AUTO, or OZAKI
 * for the residue tile, then use the recorded choice (MWGEMM_BATCH_RESIDUE_VARIANT=0|1|2
 * overrides the residue tile).  Decisions persist across processes through the file named by
 * MWGEMM_TUNE_FILE (appended on every record, read on first use).  *variant / *choice
 * receive the winner (batch: path * 16 + residue-tile variant), *best_ms its time, and
 * all_ms (optional, all_cap entries) every variant's time.
 * mwgemm_autotune_lookup returns the recorded variant for a kind (0 plain real, 1 plain
 * complex, 2 Ozaki residue GEMM, 3 batch) and shape (a, b, c), or -1; mwgemm_autotune_clear
 * forgets all decisions.  Return 0 or an error code. */
MWGEMM_API int mwgemm_autotune(int64_t m, int64_t n, int64_t k, int cplx, int path, int *variant,
                               double *best_ms, double *all_ms, int all_cap);
MWGEMM_API int mwgemm_batched_autotune(int64_t count, char transa, char transb, const int64_t *m,
                                       const int64_t *n, const int64_t *k, int *choice,
                                       double *best_ms);
MWGEMM_API int mwgemm_autotune_lookup(int kind, int64_t a, int64_t b, int64_t c);
MWGEMM_API int mwgemm_autotune_clear(void);

/* Ozaki path parameters.  mwgemm_ozaki_config reports, for contraction depth k (real or
 * complex), how many moduli the exact-product path provides for and the integer width b
 * every scaled row/column is rounded to (normwise error bound ~4*2^-b).  mwgemm_ozaki_set
 * sets the width (0 = automatic; also MWGEMM_OZAKI_BITS) and optionally forces the modulus
 * count (0 = automatic; also MWGEMM_OZAKI_NMOD).  mwgemm_ozaki_moduli lists the first nmod
 * moduli and their CRT weights y_t = ((M/m_t)^-1 mod m_t) for that set.
 *
 * Data-dependent modulus count of real products (mwgemm_ozaki_set_estimate or
 * MWGEMM_OZAKI_ESTIMATE: 1 = products of at least 2^33 multiply-adds, the default; 2 = every
 * product; 0 = never; a forced count overrides it): the width is fixed (102 unless set), the
 * residues are formed for the worst-case count, and on the device the call then keeps only
 * as many moduli as a Cauchy-Schwarz bound of its scaled operands' row norms requires.
 * Without it, the count covers the worst case k * 2^(2b) and b is raised as far as that
 * count allows (103 by default).  For real k, mwgemm_ozaki_config reports the estimating
 * plan (width and worst-case count) whenever estimation is enabled.
 * mwgemm_ozaki_last_moduli waits for the current device and reports the count the last
 * estimating call on it kept (0: none yet).  Return 0 or -1. */
MWGEMM_API int mwgemm_ozaki_config(int64_t k, int cplx, int *nmod, int *bits);
MWGEMM_API int mwgemm_ozaki_set(int bits, int nmod);
MWGEMM_API int mwgemm_ozaki_moduli(int nmod, int *m, int *y);
MWGEMM_API int mwgemm_ozaki_set_estimate(int mode);
MWGEMM_API int mwgemm_ozaki_last_moduli(int *nmod);

/* Test hook: while bytes > 0, every device buffer the library would allocate for itself
 * (exact-product workspace and modulus table, host-operand staging, batched plan memory)
 * larger than bytes is treated as unobtainable, as on a device with no more free memory, so
 * the allocation-failure behaviour can be tested with ordinary shapes; 0 lifts the limit.
 * Returns 0, or -1 for a negative size. */
MWGEMM_API int mwgemm_testing_set_alloc_limit(int64_t bytes);

/* Test hook: offset added to every data-dependent modulus count (clamped to [1, the
 * worst-case count]); 0 restores normal operation.  A negative offset deliberately keeps
 * too few moduli, so a test can show that the chosen count is needed.  Returns 0. */
MWGEMM_API int mwgemm_testing_ozaki_count_offset(int offset);

/* Test hook: runs the production INT8 MFMA residue kernel once on device buffers,
 * U[j*mp + i] = (sum_k A[i*kp + k] * B[j*kp + k]) mod m with m the mod_index-th
 * modulus of the full 40-modulus table.  mp, np multiples of 256,
 * kp a multiple of 64.  Returns 0 or an error code. */
MWGEMM_API int mwgemm_testing_residue_gemm(const void *dA, const void *dB, void *dU, int64_t mp,
                                           int64_t np, int64_t kp, int mod_index, int variant,
                                           void *stream);

/* Kernel timers: while enabled, every kernel the library launches is bracketed by HIP
 * events on its own stream.  mwgemm_timer_read synchronises, then writes up to cap
 * entries of {total milliseconds, launch count} per kernel class into ms/count and their
 * names into names (NUL-separated, names_len bytes); it returns the number of classes. */
MWGEMM_API int mwgemm_timer_enable(int on);
MWGEMM_API int mwgemm_timer_reset(void);
MWGEMM_API int mwgemm_timer_read(double *ms, int64_t *count, int cap, char *names, size_t names_len);

#ifdef __cplusplus
}
#endif

#endif /* MWGEMM_HIP_H */
