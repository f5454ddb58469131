/*
 * ddgemm_port.c -- TEST INFRASTRUCTURE ONLY (the oracle).  Not part of the
 * product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load this code, and only as the checker / the timed CPU reference.
 *
 * A from-scratch CPU restatement of the reference double-double GEMMs
 *   mgemm_  (real, float64x2)      src/epblas-parallel/multifloats/mgemm_parallel.cpp:44-191
 *   wgemm_  (complex, complex64x2) src/epblas-parallel/multifloats/wgemm_parallel.cpp:44-198
 * reproducing, per output element, the exact sequence of IEEE binary64 operations the
 * reference AVX2 path executes:
 *
 *   C := beta-prepass(C)                                   mgemm_parallel.cpp:68-77
 *   for each K chunk [pc, pc+128) in ascending order       mgemm_serial.cpp:35-37, :126
 *       acc := (0,0)                                       mgemm_serial.cpp:241-248
 *       for p in chunk ascending: acc := fast_add(acc, fast_mul(a_ip, b_pj))
 *                                                          mgemm_serial.cpp:249-269
 *       C := faithful_add(C, fast_mul(alpha, acc))         mgemm_serial.cpp:173-191
 *
 * The DD primitives follow mf_simd_fast.h:49-144 (fast set) and mf_simd_exact.h:60-99
 * (the documented op sequence of the third-party multifloats scalar operators).
 *
 * Performance shape (so the CPU baseline leg is representative of the reference):
 * GotoBLAS blocking MC/KC/NC = 64/128/256, B packed into 4-column SoA hi/lo panels, A
 * packed per 64-row block, a 4x4 AVX2 register tile, OpenMP static split over the
 * 256-column N panels (mgemm_parallel.cpp:92-137).  A scalar build of the same per-element
 * sequence runs on hosts without AVX2+FMA; both give identical bits.
 *
 * Must be compiled with -ffp-contract=off (every fused multiply-add is explicit).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <immintrin.h>

typedef struct { double hi, lo; } dd;
typedef struct { dd re, im; } zdd;

enum { PORT_MC = 64, PORT_KC = 128, PORT_NC = 256, PORT_W = 4, PORT_MR = 4 };

/* ---------------------------------------------------------------------------------
 * Scalar DD primitives.
 * ------------------------------------------------------------------------------- */

/* TwoProd by FMA: p + e == a*b exactly (mf_simd_fast.h:49-55). */
static inline void eft_prod(double a, double b, double *p, double *e)
{
    *p = a * b;
    *e = fma(a, b, -*p);
}

/* Knuth TwoSum, 6 ops, no magnitude precondition (mf_simd_fast.h:57-66). */
static inline void eft_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double bv = t - a;
    double av = t - bv;
    *e = (a - av) + (b - bv);
    *s = t;
}

/* Dekker Fast2Sum, requires |a| >= |b| (mf_simd_fast.h:68-75). */
static inline void eft_fast_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    *e = b - (t - a);
    *s = t;
}

/* Fast DD*DD (mf_simd_fast.h:79-89): TwoProd of the high limbs, both cross terms fused
 * into the error by FMA (ah*bl first, then al*bh), Fast2Sum renormalisation. */
static inline dd dd_mul_fast(dd a, dd b)
{
    double p, e;
    eft_prod(a.hi, b.hi, &p, &e);
    e = fma(a.hi, b.lo, e);
    e = fma(a.lo, b.hi, e);
    dd r;
    eft_fast_sum(p, e, &r.hi, &r.lo);
    return r;
}

/* Fast ("sloppy") DD+DD (mf_simd_fast.h:93-102). */
static inline dd dd_add_fast(dd a, dd b)
{
    double s, e;
    eft_sum(a.hi, b.hi, &s, &e);
    e = e + (a.lo + b.lo);
    dd r;
    eft_fast_sum(s, e, &r.hi, &r.lo);
    return r;
}

/* Faithful DD*DD = the multifloats scalar operator* sequence (mf_simd_exact.h:60-71):
 * cross terms as separate products, added unfused. */
static inline dd dd_mul_exact(dd a, dd b)
{
    double p, e;
    eft_prod(a.hi, b.hi, &p, &e);
    double c1 = a.hi * b.lo;
    double c2 = a.lo * b.hi;
    e = e + (c1 + c2);
    dd r;
    eft_fast_sum(p, e, &r.hi, &r.lo);
    return r;
}

/* Faithful DD+DD = the multifloats scalar operator+ sequence (mf_simd_exact.h:73-86). */
static inline dd dd_add_exact(dd a, dd b)
{
    double s, e, t, f;
    eft_sum(a.hi, b.hi, &s, &e);
    eft_sum(a.lo, b.lo, &t, &f);
    eft_fast_sum(s, t, &s, &t);
    e = e + f;
    e = e + t;
    dd r;
    eft_fast_sum(s, e, &r.hi, &r.lo);
    return r;
}

/* DD - DD as the add of the limb-negated subtrahend (mf_simd_exact.h:88-97). */
static inline dd dd_neg(dd a) { dd r = { -a.hi, -a.lo }; return r; }
static inline dd dd_sub_exact(dd a, dd b) { return dd_add_exact(a, dd_neg(b)); }

/* Complex fast product (mf_simd_fast.h:119-134) and sum (:137-144). */
static inline zdd zdd_mul_fast(zdd a, zdd b)
{
    zdd r;
    r.re = dd_add_fast(dd_mul_fast(a.re, b.re), dd_neg(dd_mul_fast(a.im, b.im)));
    r.im = dd_add_fast(dd_mul_fast(a.re, b.im), dd_mul_fast(a.im, b.re));
    return r;
}
static inline zdd zdd_add_fast(zdd a, zdd b)
{
    zdd r = { dd_add_fast(a.re, b.re), dd_add_fast(a.im, b.im) };
    return r;
}

/* Complex faithful product / sum / conjugate (mf_kernels.h:43-50). */
static inline zdd zdd_mul_exact(zdd a, zdd b)
{
    zdd r;
    r.re = dd_sub_exact(dd_mul_exact(a.re, b.re), dd_mul_exact(a.im, b.im));
    r.im = dd_add_exact(dd_mul_exact(a.re, b.im), dd_mul_exact(a.im, b.re));
    return r;
}
static inline zdd zdd_add_exact(zdd a, zdd b)
{
    zdd r = { dd_add_exact(a.re, b.re), dd_add_exact(a.im, b.im) };
    return r;
}
static inline zdd zdd_conj(zdd a) { zdd r = { a.re, dd_neg(a.im) }; return r; }

/* alpha/beta tests by limbs (mf_pred.h:28-49). */
static inline int dd_is0(dd x) { return x.hi == 0.0 && x.lo == 0.0; }
static inline int dd_is1(dd x) { return x.hi == 1.0 && x.lo == 0.0; }
static inline int zdd_is0(zdd x) { return dd_is0(x.re) && dd_is0(x.im); }
static inline int zdd_is1(zdd x) { return dd_is1(x.re) && dd_is0(x.im); }

/* Trans letters (common/blas_char.h:33-39): upper-case; real folds 'C' to 'T'.  Letters
 * outside N/T/C are not rejected by the reference; each packer reads them its own way:
 *   real: anything but 'N' is transposed (mgemm_serial.cpp:52, 71);
 *   complex op(A): 'N', 'T', anything else conjugate-transposed (wgemm_serial.cpp:56-73);
 *   complex op(B), SIMD packer: transposed, conjugated only for 'C' (wgemm_serial.cpp:145,
 *     174-185); scalar packer: 'N', 'T', anything else conjugate-transposed (:81-99);
 *   triangular-output diagonal: transposed unless 'N', conjugated only for 'C'
 *     (wgemmtr_serial.cpp:54-57). */
static inline char trans_real(char c) { c = (char)toupper((unsigned char)c); return c == 'C' ? 'T' : c; }
static inline char trans_cplx(char c) { return (char)toupper((unsigned char)c); }
static inline int conj_by_gemm_packer(char t) { return t != 'N' && t != 'T'; }

/* Exported primitive wrappers so tests can check the primitives one by one. */
void port_dd_mul_fast(const dd *a, const dd *b, dd *r) { *r = dd_mul_fast(*a, *b); }
void port_dd_add_fast(const dd *a, const dd *b, dd *r) { *r = dd_add_fast(*a, *b); }
void port_dd_mul_exact(const dd *a, const dd *b, dd *r) { *r = dd_mul_exact(*a, *b); }
void port_dd_add_exact(const dd *a, const dd *b, dd *r) { *r = dd_add_exact(*a, *b); }
void port_eft_prod(double a, double b, double *p, double *e) { eft_prod(a, b, p, e); }
void port_eft_sum(double a, double b, double *s, double *e) { eft_sum(a, b, s, e); }

/* ---------------------------------------------------------------------------------
 * Real GEMM.
 * ------------------------------------------------------------------------------- */

/* op(A)[ic:ic+ib, pc:pc+pb] -> Ap[p*ib + i] (mgemm_serial.cpp:48-65). */
static void rpack_a(const dd *A, ptrdiff_t lda, char ta, ptrdiff_t ic, ptrdiff_t pc,
                    ptrdiff_t ib, ptrdiff_t pb, dd *Ap)
{
    for (ptrdiff_t p = 0; p < pb; ++p)
        for (ptrdiff_t i = 0; i < ib; ++i)
            Ap[p * ib + i] = (ta == 'N') ? A[(size_t)(pc + p) * lda + ic + i]
                                         : A[(size_t)(ic + i) * lda + pc + p];
}

/* op(B)[pc:pc+pb, jc:jc+jb] -> SoA panels of PORT_W columns, zero padded
 * (mgemm_serial.cpp:103-158): Bh[panel*pb*W + p*W + c]. */
static void rpack_b(const dd *B, ptrdiff_t ldb, char tb, ptrdiff_t pc, ptrdiff_t jc,
                    ptrdiff_t pb, ptrdiff_t jb, double *Bh, double *Bl)
{
    ptrdiff_t np = (jb + PORT_W - 1) / PORT_W;
    for (ptrdiff_t panel = 0; panel < np; ++panel)
        for (ptrdiff_t p = 0; p < pb; ++p)
            for (ptrdiff_t c = 0; c < PORT_W; ++c) {
                ptrdiff_t j = panel * PORT_W + c;
                size_t o = (size_t)panel * pb * PORT_W + (size_t)p * PORT_W + c;
                if (j < jb) {
                    dd v = (tb == 'N') ? B[(size_t)(jc + j) * ldb + pc + p]
                                       : B[(size_t)(pc + p) * ldb + jc + j];
                    Bh[o] = v.hi; Bl[o] = v.lo;
                } else {
                    Bh[o] = 0.0; Bl[o] = 0.0;
                }
            }
}

/* Flush one accumulator into C: C := faithful_add(C, fast_mul(alpha, acc)). */
static inline void flush_real(dd *cij, dd alpha, dd acc)
{
    *cij = dd_add_exact(*cij, dd_mul_fast(alpha, acc));
}

/* Scalar micro-kernel over one (ib x jb x pb) block; same per-element sequence as the
 * vector one below. */
static void rblock_scalar(ptrdiff_t ib, ptrdiff_t jb, ptrdiff_t pb, dd alpha, const dd *Ap,
                          const double *Bh, const double *Bl, dd *C, ptrdiff_t ldc)
{
    for (ptrdiff_t j = 0; j < jb; ++j) {
        const double *bh = Bh + (size_t)(j / PORT_W) * pb * PORT_W + (j % PORT_W);
        const double *bl = Bl + (size_t)(j / PORT_W) * pb * PORT_W + (j % PORT_W);
        for (ptrdiff_t i = 0; i < ib; ++i) {
            dd acc = { 0.0, 0.0 };
            for (ptrdiff_t p = 0; p < pb; ++p) {
                dd b = { bh[p * PORT_W], bl[p * PORT_W] };
                acc = dd_add_fast(acc, dd_mul_fast(Ap[p * ib + i], b));
            }
            flush_real(&C[(size_t)j * ldc + i], alpha, acc);
        }
    }
}

/* AVX2 4-wide versions of the fast primitives: lane-wise identical to the scalar ones. */
#define V4 __attribute__((target("avx2,fma"), always_inline)) static inline
V4 void v_mul_fast(__m256d ah, __m256d al, __m256d bh, __m256d bl, __m256d *rh, __m256d *rl)
{
    __m256d p = _mm256_mul_pd(ah, bh);
    __m256d e = _mm256_fmsub_pd(ah, bh, p);
    e = _mm256_fmadd_pd(ah, bl, e);
    e = _mm256_fmadd_pd(al, bh, e);
    __m256d s = _mm256_add_pd(p, e);
    *rl = _mm256_sub_pd(e, _mm256_sub_pd(s, p));
    *rh = s;
}
V4 void v_add_fast(__m256d ah, __m256d al, __m256d bh, __m256d bl, __m256d *rh, __m256d *rl)
{
    __m256d s = _mm256_add_pd(ah, bh);
    __m256d bv = _mm256_sub_pd(s, ah);
    __m256d av = _mm256_sub_pd(s, bv);
    __m256d e = _mm256_add_pd(_mm256_sub_pd(ah, av), _mm256_sub_pd(bh, bv));
    e = _mm256_add_pd(e, _mm256_add_pd(al, bl));
    __m256d t = _mm256_add_pd(s, e);
    *rl = _mm256_sub_pd(e, _mm256_sub_pd(t, s));
    *rh = t;
}

__attribute__((target("avx2,fma")))
static void rblock_avx2(ptrdiff_t ib, ptrdiff_t jb, ptrdiff_t pb, dd alpha, const dd *Ap,
                        const double *Bh, const double *Bl, dd *C, ptrdiff_t ldc)
{
    ptrdiff_t np = (jb + PORT_W - 1) / PORT_W;
    for (ptrdiff_t panel = 0; panel < np; ++panel) {
        const double *bh = Bh + (size_t)panel * pb * PORT_W;
        const double *bl = Bl + (size_t)panel * pb * PORT_W;
        ptrdiff_t jw = jb - panel * PORT_W;
        if (jw > PORT_W) jw = PORT_W;
        dd *Cp = C + (size_t)panel * PORT_W * ldc;
        ptrdiff_t i = 0;
        for (; i < ib; i += PORT_MR) {
            ptrdiff_t rows = (ib - i < PORT_MR) ? ib - i : PORT_MR;
            __m256d ah_acc[PORT_MR], al_acc[PORT_MR];
            for (int r = 0; r < PORT_MR; ++r) { ah_acc[r] = _mm256_setzero_pd(); al_acc[r] = _mm256_setzero_pd(); }
            if (rows == PORT_MR) {
                for (ptrdiff_t p = 0; p < pb; ++p) {
                    __m256d vbh = _mm256_loadu_pd(bh + p * PORT_W);
                    __m256d vbl = _mm256_loadu_pd(bl + p * PORT_W);
                    const dd *ap = Ap + p * ib + i;
                    for (int r = 0; r < PORT_MR; ++r) {
                        __m256d mh, ml;
                        v_mul_fast(_mm256_set1_pd(ap[r].hi), _mm256_set1_pd(ap[r].lo), vbh, vbl, &mh, &ml);
                        v_add_fast(ah_acc[r], al_acc[r], mh, ml, &ah_acc[r], &al_acc[r]);
                    }
                }
            } else {
                for (ptrdiff_t p = 0; p < pb; ++p) {
                    __m256d vbh = _mm256_loadu_pd(bh + p * PORT_W);
                    __m256d vbl = _mm256_loadu_pd(bl + p * PORT_W);
                    const dd *ap = Ap + p * ib + i;
                    for (ptrdiff_t r = 0; r < rows; ++r) {
                        __m256d mh, ml;
                        v_mul_fast(_mm256_set1_pd(ap[r].hi), _mm256_set1_pd(ap[r].lo), vbh, vbl, &mh, &ml);
                        v_add_fast(ah_acc[r], al_acc[r], mh, ml, &ah_acc[r], &al_acc[r]);
                    }
                }
            }
            for (ptrdiff_t r = 0; r < rows; ++r) {
                double hs[PORT_W], ls[PORT_W];
                _mm256_storeu_pd(hs, ah_acc[r]);
                _mm256_storeu_pd(ls, al_acc[r]);
                for (ptrdiff_t c = 0; c < jw; ++c) {
                    dd acc = { hs[c], ls[c] };
                    flush_real(&Cp[(size_t)c * ldc + i + r], alpha, acc);
                }
            }
        }
    }
}

static int g_force_scalar = 0;

/* Operation order of the reference's plain kernels: 0 = its AVX2/FMA kernels (the blocked
 * code below), 1 = the scalar kernels it runs on hosts without AVX2/FMA
 * (mgemm_serial.cpp:87-99, wgemm_serial.cpp:102-115, mgemv.cpp:147-153, :213-223,
 * wgemv.cpp:170-188, :284-295), restated by the direct loops further down. */
static int g_order = 0;
void port_set_order(int order) { g_order = order; }
int port_get_order(void) { return g_order; }

/* The scalar kernels' per-element sequence: for column j and p ascending, t = alpha*b_pj
 * (faithful), then C_ij := C_ij + t*a_ip (faithful) for every row i.  The reference runs
 * this inside its GotoBLAS blocking (jc, pc, ic), which leaves each element's p order
 * ascending, so a direct loop over p is the same sequence. */
static void rgemm_scalar_order(char ta, char tb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                               dd alpha, const dd *a, ptrdiff_t lda, const dd *b,
                               ptrdiff_t ldb, dd *c, ptrdiff_t ldc, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (ptrdiff_t j = 0; j < n; ++j) {
        dd *cj = c + (size_t)j * ldc;
        for (ptrdiff_t p = 0; p < k; ++p) {
            const dd bpj = (tb == 'N') ? b[(size_t)j * ldb + p] : b[(size_t)p * ldb + j];
            const dd t = dd_mul_exact(alpha, bpj);
            for (ptrdiff_t i = 0; i < m; ++i) {
                const dd aip = (ta == 'N') ? a[(size_t)p * lda + i] : a[(size_t)i * lda + p];
                cj[i] = dd_add_exact(cj[i], dd_mul_exact(t, aip));
            }
        }
    }
}

static int have_avx2_fma(void)
{
    static int cached = -1;
    if (cached < 0) {
        __builtin_cpu_init();
        cached = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    }
    return cached && !g_force_scalar;
}

/* Select the scalar micro-kernel even on AVX2 hosts (the two must agree bit for bit). */
void port_force_scalar(int on) { g_force_scalar = on; }

/* C := alpha*op(A)*op(B) + beta*C over DD, column-major, reference mgemm_ semantics.
 * nthreads <= 0 means the OpenMP default.  Returns 0, or -1 when a pack buffer could not
 * be allocated (C is then left at beta*C and the caller is told, never silently). */
int port_mgemm(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
               const dd *alpha_, const dd *a, ptrdiff_t lda, const dd *b, ptrdiff_t ldb,
               const dd *beta_, dd *c, ptrdiff_t ldc, int nthreads)
{
    const dd alpha = *alpha_, beta = *beta_;
    const char ta = trans_real(transa), tb = trans_real(transb);
    if (m <= 0 || n <= 0) return 0;
    for (ptrdiff_t j = 0; j < n; ++j)
        for (ptrdiff_t i = 0; i < m; ++i) {
            dd *cij = &c[(size_t)j * ldc + i];
            if (dd_is0(beta)) { cij->hi = 0.0; cij->lo = 0.0; }
            else if (!dd_is1(beta)) *cij = dd_mul_exact(*cij, beta);
        }
    if (dd_is0(alpha) || k == 0) return 0;
    if (g_order == 1) {
        rgemm_scalar_order(ta, tb, m, n, k, alpha, a, lda, b, ldb, c, ldc, nthreads);
        return 0;
    }

    const int vec = have_avx2_fma();
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(|: failed)
#endif
    {
        dd *Ap = (dd *)malloc(sizeof(dd) * PORT_MC * PORT_KC);
        double *Bh = (double *)malloc(sizeof(double) * PORT_KC * PORT_NC);
        double *Bl = (double *)malloc(sizeof(double) * PORT_KC * PORT_NC);
        int ok = Ap && Bh && Bl;
        if (!ok) failed = 1;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (ptrdiff_t jc = 0; jc < n; jc += PORT_NC) {
            if (!ok) continue;
            ptrdiff_t jb = (n - jc < PORT_NC) ? n - jc : PORT_NC;
            for (ptrdiff_t pc = 0; pc < k; pc += PORT_KC) {
                ptrdiff_t pb = (k - pc < PORT_KC) ? k - pc : PORT_KC;
                rpack_b(b, ldb, tb, pc, jc, pb, jb, Bh, Bl);
                for (ptrdiff_t ic = 0; ic < m; ic += PORT_MC) {
                    ptrdiff_t ib = (m - ic < PORT_MC) ? m - ic : PORT_MC;
                    rpack_a(a, lda, ta, ic, pc, ib, pb, Ap);
                    dd *Cb = &c[(size_t)jc * ldc + ic];
                    if (vec) rblock_avx2(ib, jb, pb, alpha, Ap, Bh, Bl, Cb, ldc);
                    else     rblock_scalar(ib, jb, pb, alpha, Ap, Bh, Bl, Cb, ldc);
                }
            }
        }
        free(Ap); free(Bh); free(Bl);
    }
    return failed ? -1 : 0;
}

/* ---------------------------------------------------------------------------------
 * Complex GEMM.
 * ------------------------------------------------------------------------------- */

/* op(A) block with conjugation folded in (wgemm_serial.cpp:52-75). */
static void zpack_a(const zdd *A, ptrdiff_t lda, char ta, ptrdiff_t ic, ptrdiff_t pc,
                    ptrdiff_t ib, ptrdiff_t pb, zdd *Ap)
{
    for (ptrdiff_t p = 0; p < pb; ++p)
        for (ptrdiff_t i = 0; i < ib; ++i) {
            zdd v = (ta == 'N') ? A[(size_t)(pc + p) * lda + ic + i]
                                : A[(size_t)(ic + i) * lda + pc + p];
            Ap[p * ib + i] = conj_by_gemm_packer(ta) ? zdd_conj(v) : v;
        }
}

/* op(B) block, AoS, conjugation folded in (wgemm_serial.cpp:126-198 negates both
 * imaginary limbs for 'C'). */
static void zpack_b(const zdd *B, ptrdiff_t ldb, char tb, ptrdiff_t pc, ptrdiff_t jc,
                    ptrdiff_t pb, ptrdiff_t jb, zdd *Bp)
{
    for (ptrdiff_t j = 0; j < jb; ++j)
        for (ptrdiff_t p = 0; p < pb; ++p) {
            zdd v = (tb == 'N') ? B[(size_t)(jc + j) * ldb + pc + p]
                                : B[(size_t)(pc + p) * ldb + jc + j];
            Bp[(size_t)j * pb + p] = (tb == 'C') ? zdd_conj(v) : v;
        }
}

/* Complex block: per element acc := cadd_fast(acc, cmul_fast(a, b)) over the chunk, then
 * C := cadd_exact(C, cmul_fast(alpha, acc)) (wgemm_serial.cpp:207-356). */
static void zblock(ptrdiff_t ib, ptrdiff_t jb, ptrdiff_t pb, zdd alpha, const zdd *Ap,
                   const zdd *Bp, zdd *C, ptrdiff_t ldc)
{
    for (ptrdiff_t j = 0; j < jb; ++j) {
        const zdd *bj = Bp + (size_t)j * pb;
        for (ptrdiff_t i = 0; i < ib; ++i) {
            zdd acc = { { 0.0, 0.0 }, { 0.0, 0.0 } };
            for (ptrdiff_t p = 0; p < pb; ++p)
                acc = zdd_add_fast(acc, zdd_mul_fast(Ap[p * ib + i], bj[p]));
            zdd *cij = &C[(size_t)j * ldc + i];
            *cij = zdd_add_exact(*cij, zdd_mul_fast(alpha, acc));
        }
    }
}

int port_wgemm(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
               const zdd *alpha_, const zdd *a, ptrdiff_t lda, const zdd *b, ptrdiff_t ldb,
               const zdd *beta_, zdd *c, ptrdiff_t ldc, int nthreads)
{
    const zdd alpha = *alpha_, beta = *beta_;
    const char ta = trans_cplx(transa), tb = trans_cplx(transb);
    if (m <= 0 || n <= 0) return 0;
    for (ptrdiff_t j = 0; j < n; ++j)
        for (ptrdiff_t i = 0; i < m; ++i) {
            zdd *cij = &c[(size_t)j * ldc + i];
            if (zdd_is0(beta)) memset(cij, 0, sizeof(zdd));
            else if (!zdd_is1(beta)) *cij = zdd_mul_exact(*cij, beta);
        }
    if (zdd_is0(alpha) || k == 0) return 0;
    if (g_order == 1) {
        /* scalar kernels (wgemm_serial.cpp:77-115): t = cmul(alpha, op(B)_pj), then
         * C_ij := cadd(C_ij, cmul(t, op(A)_ip)), p ascending; 'C' operands conjugated */
#ifdef _OPENMP
        if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
        for (ptrdiff_t j = 0; j < n; ++j) {
            zdd *cj = c + (size_t)j * ldc;
            for (ptrdiff_t p = 0; p < k; ++p) {
                zdd bpj = (tb == 'N') ? b[(size_t)j * ldb + p] : b[(size_t)p * ldb + j];
                if (conj_by_gemm_packer(tb)) bpj = zdd_conj(bpj);
                const zdd t = zdd_mul_exact(alpha, bpj);
                for (ptrdiff_t i = 0; i < m; ++i) {
                    zdd aip = (ta == 'N') ? a[(size_t)p * lda + i] : a[(size_t)i * lda + p];
                    if (conj_by_gemm_packer(ta)) aip = zdd_conj(aip);
                    cj[i] = zdd_add_exact(cj[i], zdd_mul_exact(t, aip));
                }
            }
        }
        return 0;
    }

    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(|: failed)
#endif
    {
        zdd *Ap = (zdd *)malloc(sizeof(zdd) * PORT_MC * PORT_KC);
        zdd *Bp = (zdd *)malloc(sizeof(zdd) * PORT_KC * PORT_NC);
        int ok = Ap && Bp;
        if (!ok) failed = 1;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (ptrdiff_t jc = 0; jc < n; jc += PORT_NC) {
            if (!ok) continue;
            ptrdiff_t jb = (n - jc < PORT_NC) ? n - jc : PORT_NC;
            for (ptrdiff_t pc = 0; pc < k; pc += PORT_KC) {
                ptrdiff_t pb = (k - pc < PORT_KC) ? k - pc : PORT_KC;
                zpack_b(b, ldb, tb, pc, jc, pb, jb, Bp);
                for (ptrdiff_t ic = 0; ic < m; ic += PORT_MC) {
                    ptrdiff_t ib = (m - ic < PORT_MC) ? m - ic : PORT_MC;
                    zpack_a(a, lda, ta, ic, pc, ib, pb, Ap);
                    zblock(ib, jb, pb, alpha, Ap, Bp, &c[(size_t)jc * ldc + ic], ldc);
                }
            }
        }
        free(Ap); free(Bp);
    }
    return failed ? -1 : 0;
}

/* Host properties the CPU baseline reports. */
int port_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int port_has_avx2(void) { return have_avx2_fma(); }

/* ---------------------------------------------------------------------------------
 * Triangular-output GEMM: C := alpha*op(A)*op(B) + beta*C on the UPLO triangle of the
 * n x n C only (mgemmtr_ / wgemmtr_).
 *
 * Reference: src/epblas-parallel/multifloats/mgemmtr_serial.cpp:117-200 (block cores,
 * serial entry) driven by mgemmtr_parallel.cpp:27-77; complex wgemmtr_serial.cpp.  Columns
 * are walked in blocks of nb (8 unless MGEMMTR_NB / WGEMMTR_NB say otherwise,
 * mgemmtr_serial.cpp:104-112).  Per block [jc, jc+jb):
 *   - the triangle slice of each column is beta-scaled (0 -> exact zero, 1 -> kept,
 *     else C*beta with the faithful product)                    :123-130
 *   - the jb x jb diagonal triangle is updated with the scalar faithful operators,
 *     op(A) = 'N': per l, t = alpha*b_lj (skipped when t == 0), c_ij += t*a_il;
 *     otherwise:   s = sum_l a_li*b_lj (faithful, ascending l), c_ij += alpha*s   :48-95
 *   - the off-diagonal rectangle (rows above the block for 'U', below it for 'L') goes
 *     through the serial GEMM with beta = 1                     :136-166
 * The blocks touch disjoint columns; the reference's OpenMP split over blocks gives the
 * same bits as the serial sweep, and so does this one.
 * ------------------------------------------------------------------------------- */
static void rtr_diag(ptrdiff_t jc, ptrdiff_t jb, ptrdiff_t k, dd alpha, const dd *a,
                     ptrdiff_t lda, const dd *b, ptrdiff_t ldb, dd *c, ptrdiff_t ldc, int upper,
                     char ta, char tb)
{
    for (ptrdiff_t j = jc; j < jc + jb; ++j) {
        const ptrdiff_t is = upper ? jc : j, ie = upper ? j + 1 : jc + jb;
        dd *cj = c + (size_t)j * ldc;
        if (ta == 'N') {
            for (ptrdiff_t l = 0; l < k; ++l) {
                const dd bl = (tb == 'N') ? b[(size_t)j * ldb + l] : b[(size_t)l * ldb + j];
                const dd t = dd_mul_exact(alpha, bl);
                if (dd_is0(t)) continue;
                const dd *al = a + (size_t)l * lda;
                for (ptrdiff_t i = is; i < ie; ++i) cj[i] = dd_add_exact(cj[i], dd_mul_exact(t, al[i]));
            }
        } else {
            for (ptrdiff_t i = is; i < ie; ++i) {
                dd s = { 0.0, 0.0 };
                for (ptrdiff_t l = 0; l < k; ++l) {
                    const dd bl = (tb == 'N') ? b[(size_t)j * ldb + l] : b[(size_t)l * ldb + j];
                    s = dd_add_exact(s, dd_mul_exact(a[(size_t)i * lda + l], bl));
                }
                cj[i] = dd_add_exact(cj[i], dd_mul_exact(alpha, s));
            }
        }
    }
}

static void rtr_beta_cols(ptrdiff_t j0, ptrdiff_t j1, ptrdiff_t n, int upper, dd beta, dd *c,
                          ptrdiff_t ldc)
{
    for (ptrdiff_t j = j0; j < j1; ++j) {
        const ptrdiff_t is = upper ? 0 : j, ie = upper ? j + 1 : n;
        dd *cj = c + (size_t)j * ldc;
        for (ptrdiff_t i = is; i < ie; ++i) {
            if (dd_is0(beta)) { cj[i].hi = 0.0; cj[i].lo = 0.0; }
            else if (!dd_is1(beta)) cj[i] = dd_mul_exact(cj[i], beta);
        }
    }
}

/* nb <= 0 selects the reference default of 8.  Returns 0, or -1 when a GEMM pack buffer
 * could not be allocated. */
int port_mgemmtr(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                 const dd *alpha_, const dd *a, ptrdiff_t lda, const dd *b, ptrdiff_t ldb,
                 const dd *beta_, dd *c, ptrdiff_t ldc, ptrdiff_t nb, int nthreads)
{
    const dd alpha = *alpha_, beta = *beta_, one = { 1.0, 0.0 };
    const int upper = toupper((unsigned char)uplo) == 'U';
    const char ta = trans_real(transa), tb = trans_real(transb);
    if (n <= 0) return 0;
    if (dd_is0(alpha) || k == 0) {
        if (dd_is1(beta)) return 0;
        rtr_beta_cols(0, n, n, upper, beta, c, ldc);
        return 0;
    }
    if (nb <= 0) nb = 8;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(|: failed)
#endif
    for (ptrdiff_t jc = 0; jc < n; jc += nb) {
        const ptrdiff_t jb = (n - jc < nb) ? n - jc : nb;
        rtr_beta_cols(jc, jc + jb, n, upper, beta, c, ldc);
        rtr_diag(jc, jb, k, alpha, a, lda, b, ldb, c, ldc, upper, ta, tb);
        const dd *bblk = (tb == 'N') ? b + (size_t)jc * ldb : b + jc;
        if (upper) {
            if (jc > 0 && port_mgemm(ta, tb, jc, jb, k, &alpha, a, lda, bblk, ldb, &one,
                                     c + (size_t)jc * ldc, ldc, 1) != 0)
                failed = 1;
        } else {
            const ptrdiff_t r0 = jc + jb, trailing = n - r0;
            const dd *ablk = (ta == 'N') ? a + r0 : a + (size_t)r0 * lda;
            if (trailing > 0 && port_mgemm(ta, tb, trailing, jb, k, &alpha, ablk, lda, bblk, ldb,
                                           &one, c + (size_t)jc * ldc + r0, ldc, 1) != 0)
                failed = 1;
        }
    }
    return failed ? -1 : 0;
}

static inline zdd zop(zdd x, int conj) { return conj ? zdd_conj(x) : x; }

static void ztr_diag(ptrdiff_t jc, ptrdiff_t jb, ptrdiff_t k, zdd alpha, const zdd *a,
                     ptrdiff_t lda, const zdd *b, ptrdiff_t ldb, zdd *c, ptrdiff_t ldc, int upper,
                     char ta, char tb)
{
    const int trans_a = ta != 'N', trans_b = tb != 'N', conj_a = ta == 'C', conj_b = tb == 'C';
    for (ptrdiff_t j = jc; j < jc + jb; ++j) {
        const ptrdiff_t is = upper ? jc : j, ie = upper ? j + 1 : jc + jb;
        zdd *cj = c + (size_t)j * ldc;
        if (!trans_a) {
            for (ptrdiff_t l = 0; l < k; ++l) {
                const zdd bl = trans_b ? zop(b[(size_t)l * ldb + j], conj_b) : b[(size_t)j * ldb + l];
                const zdd t = zdd_mul_exact(alpha, bl);
                if (zdd_is0(t)) continue;
                const zdd *al = a + (size_t)l * lda;
                for (ptrdiff_t i = is; i < ie; ++i) cj[i] = zdd_add_exact(cj[i], zdd_mul_exact(t, al[i]));
            }
        } else {
            for (ptrdiff_t i = is; i < ie; ++i) {
                zdd s = { { 0.0, 0.0 }, { 0.0, 0.0 } };
                for (ptrdiff_t l = 0; l < k; ++l) {
                    const zdd al = zop(a[(size_t)i * lda + l], conj_a);
                    const zdd bl = trans_b ? zop(b[(size_t)l * ldb + j], conj_b) : b[(size_t)j * ldb + l];
                    s = zdd_add_exact(s, zdd_mul_exact(al, bl));
                }
                cj[i] = zdd_add_exact(cj[i], zdd_mul_exact(alpha, s));
            }
        }
    }
}

static void ztr_beta_cols(ptrdiff_t j0, ptrdiff_t j1, ptrdiff_t n, int upper, zdd beta, zdd *c,
                          ptrdiff_t ldc)
{
    for (ptrdiff_t j = j0; j < j1; ++j) {
        const ptrdiff_t is = upper ? 0 : j, ie = upper ? j + 1 : n;
        zdd *cj = c + (size_t)j * ldc;
        for (ptrdiff_t i = is; i < ie; ++i) {
            if (zdd_is0(beta)) memset(&cj[i], 0, sizeof(zdd));
            else if (!zdd_is1(beta)) cj[i] = zdd_mul_exact(cj[i], beta);
        }
    }
}

int port_wgemmtr(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                 const zdd *alpha_, const zdd *a, ptrdiff_t lda, const zdd *b, ptrdiff_t ldb,
                 const zdd *beta_, zdd *c, ptrdiff_t ldc, ptrdiff_t nb, int nthreads)
{
    const zdd alpha = *alpha_, beta = *beta_, one = { { 1.0, 0.0 }, { 0.0, 0.0 } };
    const int upper = toupper((unsigned char)uplo) == 'U';
    const char ta = trans_cplx(transa), tb = trans_cplx(transb);
    if (n <= 0) return 0;
    if (zdd_is0(alpha) || k == 0) {
        if (zdd_is1(beta)) return 0;
        ztr_beta_cols(0, n, n, upper, beta, c, ldc);
        return 0;
    }
    if (nb <= 0) nb = 8;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(|: failed)
#endif
    for (ptrdiff_t jc = 0; jc < n; jc += nb) {
        const ptrdiff_t jb = (n - jc < nb) ? n - jc : nb;
        ztr_beta_cols(jc, jc + jb, n, upper, beta, c, ldc);
        ztr_diag(jc, jb, k, alpha, a, lda, b, ldb, c, ldc, upper, ta, tb);
        const zdd *bblk = (tb == 'N') ? b + (size_t)jc * ldb : b + jc;
        if (upper) {
            if (jc > 0 && port_wgemm(ta, tb, jc, jb, k, &alpha, a, lda, bblk, ldb, &one,
                                     c + (size_t)jc * ldc, ldc, 1) != 0)
                failed = 1;
        } else {
            const ptrdiff_t r0 = jc + jb, trailing = n - r0;
            const zdd *ablk = (ta == 'N') ? a + r0 : a + (size_t)r0 * lda;
            if (trailing > 0 && port_wgemm(ta, tb, trailing, jb, k, &alpha, ablk, lda, bblk, ldb,
                                           &one, c + (size_t)jc * ldc + r0, ldc, 1) != 0)
                failed = 1;
        }
    }
    return failed ? -1 : 0;
}

/* ---------------------------------------------------------------------------------
 * Matrix-vector products mgemv_ / wgemv_: y := alpha*op(A)*x + beta*y, reference AVX2
 * numerics (src/epblas-parallel/multifloats/mgemv.cpp:225-275, wgemv.cpp).
 *   m == 0 or n == 0: return before anything (mgemv.cpp:250); then y := beta*y
 *   (0 -> exact zero, 1 -> kept, else faithful beta*y; mf_kernels.h:63-74); alpha == 0:
 *   return.  Strided vectors are gathered/scattered, negative strides from the far end
 *   (mf_kernels.h:84-93); the arithmetic is that of the unit-stride cores:
 *   op 'N' (mgemv.cpp:60-105): rows are split into nbands bands [lo, hi)
 *     (blas_part_bound, the OpenMP team of the reference); within a band rows in whole
 *     4-blocks from lo take y := fast_add(y, fast_mul(t_j, a_ij)), the last (hi-lo) % 4
 *     rows take y := faithful_add(y, faithful_mul(t_j, a_ij)); t_j = faithful(alpha*x_j),
 *     columns with x_j == 0 skipped, j ascending.  Complex (wgemv.cpp:55-150): the band
 *     bounds are floored to multiples of 4 and the last band runs to m, so only the final
 *     m % 4 rows take the faithful tail.
 *   op 'T'/'C' (mgemv.cpp:150-205): per column four accumulators s_L over rows i = L mod 4
 *     below m - m%4 (fast ops, s := fast_add(s, fast_mul(a_ij, x_i))), reduced as
 *     fast_add(fast_add(s0, s2), fast_add(s1, s3)); the last m % 4 rows faithfully
 *     (s := s + a_ij*x_i); y_j := faithful_add(y_j, faithful_mul(alpha, s)).  'C' conjugates
 *     A in the complex case.
 * ------------------------------------------------------------------------------- */
static inline ptrdiff_t part_bound(ptrdiff_t total, ptrdiff_t idx, ptrdiff_t nparts)
{
    const ptrdiff_t base = total / nparts, rem = total % nparts;
    return idx * base + (idx < rem ? idx : rem);
}

static inline ptrdiff_t vec_base(ptrdiff_t len, ptrdiff_t inc)
{
    return inc < 0 ? -(len - 1) * inc : 0;
}

int port_mgemv(char trans, ptrdiff_t m, ptrdiff_t n, const dd *alpha_, const dd *a,
               ptrdiff_t lda, const dd *x, ptrdiff_t incx, const dd *beta_, dd *y,
               ptrdiff_t incy, ptrdiff_t nbands)
{
    const dd alpha = *alpha_, beta = *beta_;
    const int notrans = trans_real(trans) == 'N';
    if (m == 0 || n == 0) return 0;
    const ptrdiff_t leny = notrans ? m : n, lenx = notrans ? n : m;
    dd *xs = (dd *)malloc(sizeof(dd) * (size_t)lenx), *ys = (dd *)malloc(sizeof(dd) * (size_t)leny);
    if (!xs || !ys) { free(xs); free(ys); return -1; }
    const ptrdiff_t bx = vec_base(lenx, incx), by = vec_base(leny, incy);
    for (ptrdiff_t i = 0; i < lenx; ++i) xs[i] = x[bx + i * incx];
    for (ptrdiff_t i = 0; i < leny; ++i) {
        dd v = y[by + i * incy];
        if (dd_is0(beta)) { v.hi = 0.0; v.lo = 0.0; }
        else if (!dd_is1(beta)) v = dd_mul_exact(beta, v);
        ys[i] = v;
    }
    if (!dd_is0(alpha) && g_order == 1) {
        /* scalar kernels (mgemv.cpp:147-153, :213-223): every row and every column term
         * with the faithful operations, one sequential dot per column for op 'T' */
        if (notrans) {
            for (ptrdiff_t j = 0; j < n; ++j) {
                if (dd_is0(xs[j])) continue;
                const dd tj = dd_mul_exact(alpha, xs[j]);
                const dd *aj = a + (size_t)j * lda;
                for (ptrdiff_t i = 0; i < m; ++i) ys[i] = dd_add_exact(ys[i], dd_mul_exact(tj, aj[i]));
            }
        } else {
            for (ptrdiff_t j = 0; j < n; ++j) {
                const dd *aj = a + (size_t)j * lda;
                dd s = { 0.0, 0.0 };
                for (ptrdiff_t i = 0; i < m; ++i) s = dd_add_exact(s, dd_mul_exact(aj[i], xs[i]));
                ys[j] = dd_add_exact(ys[j], dd_mul_exact(alpha, s));
            }
        }
    } else if (!dd_is0(alpha)) {
        if (notrans) {
            if (nbands < 1) nbands = 1;
            for (ptrdiff_t t = 0; t < nbands; ++t) {
                const ptrdiff_t lo = part_bound(m, t, nbands), hi = part_bound(m, t + 1, nbands);
                const ptrdiff_t vend = lo + (hi - lo) / 4 * 4;
                for (ptrdiff_t j = 0; j < n; ++j) {
                    if (dd_is0(xs[j])) continue;
                    const dd tj = dd_mul_exact(alpha, xs[j]);
                    const dd *aj = a + (size_t)j * lda;
                    for (ptrdiff_t i = lo; i < vend; ++i) ys[i] = dd_add_fast(ys[i], dd_mul_fast(tj, aj[i]));
                    for (ptrdiff_t i = vend; i < hi; ++i) ys[i] = dd_add_exact(ys[i], dd_mul_exact(tj, aj[i]));
                }
            }
        } else {
            const ptrdiff_t m4 = m / 4 * 4;
            for (ptrdiff_t j = 0; j < n; ++j) {
                const dd *aj = a + (size_t)j * lda;
                dd s4[4] = { { 0, 0 }, { 0, 0 }, { 0, 0 }, { 0, 0 } };
                for (ptrdiff_t i = 0; i < m4; ++i) s4[i & 3] = dd_add_fast(s4[i & 3], dd_mul_fast(aj[i], xs[i]));
                dd s = dd_add_fast(dd_add_fast(s4[0], s4[2]), dd_add_fast(s4[1], s4[3]));
                for (ptrdiff_t i = m4; i < m; ++i) s = dd_add_exact(s, dd_mul_exact(aj[i], xs[i]));
                ys[j] = dd_add_exact(ys[j], dd_mul_exact(alpha, s));
            }
        }
    }
    for (ptrdiff_t i = 0; i < leny; ++i) y[by + i * incy] = ys[i];
    free(xs); free(ys);
    return 0;
}

int port_wgemv(char trans, ptrdiff_t m, ptrdiff_t n, const zdd *alpha_, const zdd *a,
               ptrdiff_t lda, const zdd *x, ptrdiff_t incx, const zdd *beta_, zdd *y,
               ptrdiff_t incy)
{
    const zdd alpha = *alpha_, beta = *beta_;
    const char tr = trans_cplx(trans);
    const int notrans = tr == 'N', conj_a = tr == 'C';
    if (m == 0 || n == 0) return 0;
    const ptrdiff_t leny = notrans ? m : n, lenx = notrans ? n : m;
    zdd *xs = (zdd *)malloc(sizeof(zdd) * (size_t)lenx), *ys = (zdd *)malloc(sizeof(zdd) * (size_t)leny);
    if (!xs || !ys) { free(xs); free(ys); return -1; }
    const ptrdiff_t bx = vec_base(lenx, incx), by = vec_base(leny, incy);
    for (ptrdiff_t i = 0; i < lenx; ++i) xs[i] = x[bx + i * incx];
    for (ptrdiff_t i = 0; i < leny; ++i) {
        zdd v = y[by + i * incy];
        if (zdd_is0(beta)) memset(&v, 0, sizeof v);
        else if (!zdd_is1(beta)) v = zdd_mul_exact(beta, v);
        ys[i] = v;
    }
    if (!zdd_is0(alpha) && g_order == 1) {
        /* scalar kernels (wgemv.cpp:170-188, :284-295) */
        if (notrans) {
            for (ptrdiff_t j = 0; j < n; ++j) {
                if (zdd_is0(xs[j])) continue;
                const zdd tj = zdd_mul_exact(alpha, xs[j]);
                const zdd *aj = a + (size_t)j * lda;
                for (ptrdiff_t i = 0; i < m; ++i) ys[i] = zdd_add_exact(ys[i], zdd_mul_exact(tj, aj[i]));
            }
        } else {
            for (ptrdiff_t j = 0; j < n; ++j) {
                const zdd *aj = a + (size_t)j * lda;
                zdd s;
                memset(&s, 0, sizeof s);
                for (ptrdiff_t i = 0; i < m; ++i)
                    s = zdd_add_exact(s, zdd_mul_exact(zop(aj[i], conj_a), xs[i]));
                ys[j] = zdd_add_exact(ys[j], zdd_mul_exact(alpha, s));
            }
        }
    } else if (!zdd_is0(alpha)) {
        const ptrdiff_t m4 = m / 4 * 4;
        if (notrans) {
            for (ptrdiff_t j = 0; j < n; ++j) {
                if (zdd_is0(xs[j])) continue;
                const zdd tj = zdd_mul_exact(alpha, xs[j]);
                const zdd *aj = a + (size_t)j * lda;
                for (ptrdiff_t i = 0; i < m4; ++i) ys[i] = zdd_add_fast(ys[i], zdd_mul_fast(tj, aj[i]));
                for (ptrdiff_t i = m4; i < m; ++i) ys[i] = zdd_add_exact(ys[i], zdd_mul_exact(tj, aj[i]));
            }
        } else {
            for (ptrdiff_t j = 0; j < n; ++j) {
                const zdd *aj = a + (size_t)j * lda;
                zdd s4[4];
                memset(s4, 0, sizeof s4);
                for (ptrdiff_t i = 0; i < m4; ++i)
                    s4[i & 3] = zdd_add_fast(s4[i & 3], zdd_mul_fast(zop(aj[i], conj_a), xs[i]));
                zdd s = zdd_add_fast(zdd_add_fast(s4[0], s4[2]), zdd_add_fast(s4[1], s4[3]));
                for (ptrdiff_t i = m4; i < m; ++i) s = zdd_add_exact(s, zdd_mul_exact(zop(aj[i], conj_a), xs[i]));
                ys[j] = zdd_add_exact(ys[j], zdd_mul_exact(alpha, s));
            }
        }
    }
    for (ptrdiff_t i = 0; i < leny; ++i) y[by + i * incy] = ys[i];
    free(xs); free(ys);
    return 0;
}
