/*
 * q128_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).  Loaded by tests/, the smoke
 * check and bench.py's cpu_baseline leg, never by the product.
 *
 * A from-scratch binary128 (__float128, libquadmath soft-float) GEMM used as the accuracy
 * oracle for the double-double GEMMs: the textbook triple loop
 *     C_q(i,j) = alpha_q * sum_{p ascending} a_q(i,p) * b_q(p,j) + beta_q * C0_q(i,j)
 * with every DD input converted exactly to binary128 (hi + lo spans <= 107 significant
 * bits for a normalised DD, which binary128's 113-bit significand holds exactly).
 * Accumulation error is ~K * 2^-113 relative to the operand norm, four orders of
 * magnitude below the 1e-30 acceptance bar.
 *
 * The normwise error it reports is the reference fuzz metric
 *     err = max|C - C_ref| / max(max|C_ref|, op_norm),
 *     op_norm = |alpha| * K * max|A| * max|B| + |beta| * max|C0|
 * (test/epblas-parallel/common/target_multifloats/fuzz_util.fypp:252-284,
 *  test/epblas-parallel/consistency/target_multifloats/fuzz_mgemm.fypp:88-90), evaluated
 * in binary128; |z| for complex entries is the modulus (fuzz_util.fypp:303-313).
 * The reference kind16 GEMMs (qgemm_/xgemm_, src/epblas-parallel/kind16/qgemm_parallel.c,
 * xgemm_parallel.c) are the independent cross-check of this file when they can be built.
 */
#include <quadmath.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __float128 q;
typedef struct { double hi, lo; } dd;
typedef struct { dd re, im; } zdd;
typedef struct { q re, im; } zq;

static inline q dd2q(dd x) { return (q)x.hi + (q)x.lo; }
static inline q qabs(q x) { return x < 0 ? -x : x; }
static inline q zqabs(zq z) { return sqrtq(z.re * z.re + z.im * z.im); }
static inline char up(char c) { return (char)toupper((unsigned char)c); }

/* op(A)(i,p) for a real matrix stored column-major. */
static inline dd rget(const dd *A, ptrdiff_t ld, char t, ptrdiff_t i, ptrdiff_t p)
{
    return (t == 'N') ? A[(size_t)p * ld + i] : A[(size_t)i * ld + p];
}
static inline zq zget(const zdd *A, ptrdiff_t ld, char t, ptrdiff_t i, ptrdiff_t p)
{
    zdd v = (t == 'N') ? A[(size_t)p * ld + i] : A[(size_t)i * ld + p];
    zq r = { dd2q(v.re), dd2q(v.im) };
    if (t == 'C') r.im = -r.im;
    return r;
}

/* Round a binary128 value to the nearest double-double (hi = nearest double, lo = the
 * nearest double to the remainder). */
static inline dd q2dd(q x)
{
    dd r;
    r.hi = (double)x;
    r.lo = (double)(x - (q)r.hi);
    return r;
}

/* Real oracle GEMM on a set of entries.  ent_i/ent_j (length nent) select output entries;
 * nent < 0 means all m*n entries in column-major order.  Outputs, per selected entry, the
 * binary128 result as 16 raw bytes (out_q, may be NULL) and/or rounded to DD (out_dd, may
 * be NULL). */
void q128_mgemm_entries(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                        const dd *alpha, const dd *a, ptrdiff_t lda, const dd *b, ptrdiff_t ldb,
                        const dd *beta, const dd *c0, ptrdiff_t ldc,
                        ptrdiff_t nent, const int64_t *ent_i, const int64_t *ent_j,
                        void *out_q, dd *out_dd)
{
    char ta = up(transa), tb = up(transb);
    if (ta == 'C') ta = 'T';
    if (tb == 'C') tb = 'T';
    const q al = dd2q(*alpha), be = dd2q(*beta);
    const int beta_zero = (beta->hi == 0.0 && beta->lo == 0.0);
    ptrdiff_t total = nent < 0 ? m * n : nent;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (ptrdiff_t e = 0; e < total; ++e) {
        ptrdiff_t i = nent < 0 ? e % m : (ptrdiff_t)ent_i[e];
        ptrdiff_t j = nent < 0 ? e / m : (ptrdiff_t)ent_j[e];
        q s = 0;
        for (ptrdiff_t p = 0; p < k; ++p)
            s += dd2q(rget(a, lda, ta, i, p)) * dd2q(rget(b, ldb, tb, p, j));
        q r = al * s;
        if (!beta_zero) r += be * dd2q(c0[(size_t)j * ldc + i]);
        if (out_q) memcpy((char *)out_q + 16 * e, &r, 16);
        if (out_dd) out_dd[e] = q2dd(r);
    }
}

void q128_wgemm_entries(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                        const zdd *alpha, const zdd *a, ptrdiff_t lda, const zdd *b, ptrdiff_t ldb,
                        const zdd *beta, const zdd *c0, ptrdiff_t ldc,
                        ptrdiff_t nent, const int64_t *ent_i, const int64_t *ent_j,
                        void *out_q, zdd *out_dd)
{
    const char ta = up(transa), tb = up(transb);
    const zq al = { dd2q(alpha->re), dd2q(alpha->im) };
    const zq be = { dd2q(beta->re), dd2q(beta->im) };
    const int beta_zero = (beta->re.hi == 0.0 && beta->re.lo == 0.0 &&
                           beta->im.hi == 0.0 && beta->im.lo == 0.0);
    ptrdiff_t total = nent < 0 ? m * n : nent;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (ptrdiff_t e = 0; e < total; ++e) {
        ptrdiff_t i = nent < 0 ? e % m : (ptrdiff_t)ent_i[e];
        ptrdiff_t j = nent < 0 ? e / m : (ptrdiff_t)ent_j[e];
        q sr = 0, si = 0;
        for (ptrdiff_t p = 0; p < k; ++p) {
            zq x = zget(a, lda, ta, i, p), y = zget(b, ldb, tb, p, j);
            sr += x.re * y.re - x.im * y.im;
            si += x.re * y.im + x.im * y.re;
        }
        zq r = { al.re * sr - al.im * si, al.re * si + al.im * sr };
        if (!beta_zero) {
            zdd cv = c0[(size_t)j * ldc + i];
            q cr = dd2q(cv.re), ci = dd2q(cv.im);
            r.re += be.re * cr - be.im * ci;
            r.im += be.re * ci + be.im * cr;
        }
        if (out_q) memcpy((char *)out_q + 32 * e, &r, 32);
        if (out_dd) { out_dd[e].re = q2dd(r.re); out_dd[e].im = q2dd(r.im); }
    }
}

/* max|x| over the logical rows x cols region of a column-major DD matrix. */
static q rmaxabs(const dd *A, ptrdiff_t rows, ptrdiff_t cols, ptrdiff_t ld)
{
    q mx = 0;
    for (ptrdiff_t j = 0; j < cols; ++j)
        for (ptrdiff_t i = 0; i < rows; ++i) {
            q v = qabs(dd2q(A[(size_t)j * ld + i]));
            if (v > mx) mx = v;
        }
    return mx;
}
static q zmaxabs(const zdd *A, ptrdiff_t rows, ptrdiff_t cols, ptrdiff_t ld)
{
    q mx = 0;
    for (ptrdiff_t j = 0; j < cols; ++j)
        for (ptrdiff_t i = 0; i < rows; ++i) {
            zdd v = A[(size_t)j * ld + i];
            zq z = { dd2q(v.re), dd2q(v.im) };
            q a = zqabs(z);
            if (a > mx) mx = a;
        }
    return mx;
}

/* Normwise error of a DD result against the oracle on a set of entries (nent < 0: all).
 * out[0] = err (normwise, as above), out[1] = max|diff|, out[2] = max|C_ref| over the
 * entries, out[3] = op_norm; all rounded to double. */
void q128_mgemm_err(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                    const dd *alpha, const dd *a, ptrdiff_t lda, const dd *b, ptrdiff_t ldb,
                    const dd *beta, const dd *c0, ptrdiff_t ldc,
                    const dd *ctest, ptrdiff_t ldct,
                    ptrdiff_t nent, const int64_t *ent_i, const int64_t *ent_j, double *out)
{
    char ta = up(transa), tb = up(transb);
    ptrdiff_t total = nent < 0 ? m * n : nent;
    q *ref = (q *)malloc(sizeof(q) * (size_t)(total > 0 ? total : 1));
    q128_mgemm_entries(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc,
                       nent, ent_i, ent_j, ref, NULL);
    q dmax = 0, rmax = 0;
    for (ptrdiff_t e = 0; e < total; ++e) {
        ptrdiff_t i = nent < 0 ? e % m : (ptrdiff_t)ent_i[e];
        ptrdiff_t j = nent < 0 ? e / m : (ptrdiff_t)ent_j[e];
        q d = qabs(dd2q(ctest[(size_t)j * ldct + i]) - ref[e]);
        if (d > dmax) dmax = d;
        if (qabs(ref[e]) > rmax) rmax = qabs(ref[e]);
    }
    ptrdiff_t ar = (ta == 'N') ? m : k, ac = (ta == 'N') ? k : m;
    ptrdiff_t br = (tb == 'N') ? k : n, bc = (tb == 'N') ? n : k;
    q opn = qabs(dd2q(*alpha)) * (q)k * rmaxabs(a, ar, ac, lda) * rmaxabs(b, br, bc, ldb);
    if (!(beta->hi == 0.0 && beta->lo == 0.0)) opn += qabs(dd2q(*beta)) * rmaxabs(c0, m, n, ldc);
    q den = rmax > opn ? rmax : opn;
    out[0] = (den < (q)2.2250738585072014e-308) ? (double)dmax : (double)(dmax / den);
    out[1] = (double)dmax; out[2] = (double)rmax; out[3] = (double)opn;
    free(ref);
}

void q128_wgemm_err(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                    const zdd *alpha, const zdd *a, ptrdiff_t lda, const zdd *b, ptrdiff_t ldb,
                    const zdd *beta, const zdd *c0, ptrdiff_t ldc,
                    const zdd *ctest, ptrdiff_t ldct,
                    ptrdiff_t nent, const int64_t *ent_i, const int64_t *ent_j, double *out)
{
    char ta = up(transa), tb = up(transb);
    ptrdiff_t total = nent < 0 ? m * n : nent;
    zq *ref = (zq *)malloc(sizeof(zq) * (size_t)(total > 0 ? total : 1));
    q128_wgemm_entries(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc,
                       nent, ent_i, ent_j, ref, NULL);
    q dmax = 0, rmax = 0;
    for (ptrdiff_t e = 0; e < total; ++e) {
        ptrdiff_t i = nent < 0 ? e % m : (ptrdiff_t)ent_i[e];
        ptrdiff_t j = nent < 0 ? e / m : (ptrdiff_t)ent_j[e];
        zdd t = ctest[(size_t)j * ldct + i];
        zq d = { dd2q(t.re) - ref[e].re, dd2q(t.im) - ref[e].im };
        q ad = zqabs(d), ar_ = zqabs(ref[e]);
        if (ad > dmax) dmax = ad;
        if (ar_ > rmax) rmax = ar_;
    }
    ptrdiff_t ar = (ta == 'N') ? m : k, ac = (ta == 'N') ? k : m;
    ptrdiff_t br = (tb == 'N') ? k : n, bc = (tb == 'N') ? n : k;
    zq alq = { dd2q(alpha->re), dd2q(alpha->im) }, beq = { dd2q(beta->re), dd2q(beta->im) };
    q opn = zqabs(alq) * (q)k * zmaxabs(a, ar, ac, lda) * zmaxabs(b, br, bc, ldb);
    opn += zqabs(beq) * zmaxabs(c0, m, n, ldc);
    q den = rmax > opn ? rmax : opn;
    out[0] = (den < (q)2.2250738585072014e-308) ? (double)dmax : (double)(dmax / den);
    out[1] = (double)dmax; out[2] = (double)rmax; out[3] = (double)opn;
    free(ref);
}

/* Helpers for the golden-vector checks: DD -> binary128 (16 raw bytes each), and the
 * normwise distance between two binary128 arrays relative to a given scale. */
void q128_from_dd(const dd *x, ptrdiff_t n, void *out)
{
    for (ptrdiff_t i = 0; i < n; ++i) {
        q v = dd2q(x[i]);
        memcpy((char *)out + 16 * i, &v, 16);
    }
}
double q128_max_abs_diff(const void *x, const void *y, ptrdiff_t n)
{
    q mx = 0;
    for (ptrdiff_t i = 0; i < n; ++i) {
        q a, b;
        memcpy(&a, (const char *)x + 16 * i, 16);
        memcpy(&b, (const char *)y + 16 * i, 16);
        q d = qabs(a - b);
        if (d > mx) mx = d;
    }
    return (double)mx;
}
double q128_max_abs(const void *x, ptrdiff_t n)
{
    q mx = 0;
    for (ptrdiff_t i = 0; i < n; ++i) {
        q a;
        memcpy(&a, (const char *)x + 16 * i, 16);
        if (qabs(a) > mx) mx = qabs(a);
    }
    return (double)mx;
}
/* Exact rational checks of the error-free transforms: returns 1 iff hi + lo == x*y
 * (resp. x + y) exactly, evaluated in binary128 (exact for these operand widths). */
int q128_check_prod(double x, double y, double hi, double lo)
{
    return (q)x * (q)y == (q)hi + (q)lo;
}
int q128_check_sum(double x, double y, double hi, double lo)
{
    return (q)x + (q)y == (q)hi + (q)lo;
}
/* |(hi+lo) - (x.hi+x.lo)*(y.hi+y.lo)| / |exact| in binary128 (the relative error of a DD
 * product; binary128 holds each DD exactly and the product to 2^-113 relative). */
double q128_dd_mul_relerr(const dd *x, const dd *y, const dd *r)
{
    q ex = dd2q(*x) * dd2q(*y);
    if (ex == 0) return (double)qabs(dd2q(*r));
    return (double)(qabs(dd2q(*r) - ex) / qabs(ex));
}
double q128_dd_add_abserr(const dd *x, const dd *y, const dd *r)
{
    q ex = dd2q(*x) + dd2q(*y);
    return (double)qabs(dd2q(*r) - ex);
}
