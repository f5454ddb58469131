// autotune.cpp -- the library's own autotuner.
//
// Decisions are kept per shape class: (kind, ceil(log2 a), ceil(log2 b), ceil(log2 c)) where
// (a, b, c) is (m, n, k) for single GEMMs (the Ozaki kind keys on the padded residue-GEMM
// shape) and (count, cube root of the mean m*n*k, max k) for batches.  A tuning run times
// every candidate on synthetic operands of the caller's shapes (device memory it allocates
// and frees itself, so the caller's data is never touched), records the fastest, and
// optionally appends the decision to the file named by MWGEMM_TUNE_FILE, which is read back
// on first use.  Dispatch consults the table and falls back to built-in heuristics.
// The number of moduli of the Ozaki path is not tuned: it is the smallest count that meets
// the guaranteed error bound for the depth k.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/mwgemm_hip.h"
#include "mw_internal.hpp"
#include "ozaki.hpp"

namespace mw {

double batch_time_variant(int64_t count, char ta, char tb, const int64_t *m, const int64_t *n,
                          const int64_t *k, const mw_float64x2 *alpha,
                          const mw_float64x2 *const *dA, const int64_t *lda,
                          const mw_float64x2 *const *dB, const int64_t *ldb,
                          const mw_float64x2 *beta, mw_float64x2 *const *dC, const int64_t *ldc,
                          int path, int oz_variant, int reps, hipStream_t s);

namespace {

typedef std::tuple<int, int, int, int> Key;
struct Entry { int variant; double ms; };
std::mutex g_mu;
std::map<Key, Entry> g_table;
bool g_loaded = false;

int clog2(int64_t x)
{
    int l = 0;
    while (l < 62 && ((int64_t)1 << l) < x) ++l;
    return l;
}

Key make_key(int kind, int64_t a, int64_t b, int64_t c)
{
    return Key(kind, clog2(a), clog2(b), clog2(c));
}

void load_locked()
{
    if (g_loaded) return;
    g_loaded = true;
    const char *path = getenv("MWGEMM_TUNE_FILE");
    if (!path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    int kind, a, b, c, v;
    double ms;
    while (fscanf(f, "%d %d %d %d %d %lf", &kind, &a, &b, &c, &v, &ms) == 6)
        if (kind >= 0 && kind < TK_COUNT) g_table[Key(kind, a, b, c)] = Entry{v, ms};
    fclose(f);
}

// Synthetic DD operand: hi uniform in [-1,1) from a counter hash, lo = hi * 2^-60.
__global__ __launch_bounds__(256)
void k_fill_dd(ddv *p, int64_t n, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double h = (double)(z >> 11) * 0x1p-52 - 1.0;
    p[i] = ddv{h, h * 0x1p-60};
}

int fill(ddv *p, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_dd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       (uint64_t)n * 7919u);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

int tune_lookup(int kind, int64_t a, int64_t b, int64_t c)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked();
    auto it = g_table.find(make_key(kind, a, b, c));
    return it == g_table.end() ? -1 : it->second.variant;
}

void tune_record(int kind, int64_t a, int64_t b, int64_t c, int variant, double ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked();
    const Key key = make_key(kind, a, b, c);
    g_table[key] = Entry{variant, ms};
    if (const char *path = getenv("MWGEMM_TUNE_FILE")) {
        if (FILE *f = fopen(path, "a")) {
            fprintf(f, "%d %d %d %d %d %.6f\n", std::get<0>(key), std::get<1>(key),
                    std::get<2>(key), std::get<3>(key), variant, ms);
            fclose(f);
        }
    }
}

}  // namespace mw

using namespace mw;

extern "C" {

int mwgemm_autotune(int64_t m, int64_t n, int64_t k, int cplx, int path, int *variant,
                    double *best_ms, double *all_ms, int all_cap)
{
    if (m <= 0 || n <= 0 || k <= 0 || (path != MWGEMM_PATH_PLAIN && path != MWGEMM_PATH_OZAKI))
        return record_error(-3, "autotune: bad shape or path");
    if (cplx && path == MWGEMM_PATH_PLAIN) {   // one complex plain tile: nothing to choose
        if (variant) *variant = 0;
        if (best_ms) *best_ms = 0.0;
        return 0;
    }
    const size_t esz = cplx ? sizeof(zddv) : sizeof(ddv);
    void *A = nullptr, *B = nullptr, *C = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&A, esz * m * k) != hipSuccess || hipMalloc(&B, esz * k * n) != hipSuccess ||
        hipMalloc(&C, esz * m * n) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    {
        (void)hipGetLastError();
        rc = record_error(-6, "autotune: cannot allocate %zu bytes of probe operands", esz * (m * k + k * n + m * n));
    }
    const int f = cplx ? 2 : 1;
    if (!rc) rc = fill((ddv *)A, f * m * k, s);
    if (!rc) rc = fill((ddv *)B, f * k * n, s);
    if (!rc) rc = fill((ddv *)C, f * m * n, s);
    const int nvar = path == MWGEMM_PATH_PLAIN ? PLAIN_REAL_VARIANTS : OZ_GEMM_VARIANTS;
    int best = -1;
    double best_t = 0.0;
    for (int v = 0; v < nvar && !rc; ++v) {
        double t_v = -1.0;
        for (int rep = 0; rep < 3 && !rc; ++rep) {
            int r2 = 0;
            if (hipEventRecord(e0, s) != hipSuccess) { rc = -1; break; }
            if (!cplx) {
                GemmArgsR g{};
                g.m = m; g.n = n; g.k = k; g.opa = OP_N; g.opb = OP_N;
                g.alpha = ddv{1.0, 0.0}; g.beta = ddv{0.5, 0.0};
                g.A = (const ddv *)A; g.lda = m; g.B = (const ddv *)B; g.ldb = k;
                g.C = (ddv *)C; g.ldc = m; g.variant = v;
                r2 = gemm_real_device(g, path, s);
            } else {
                GemmArgsZ g{};
                g.m = m; g.n = n; g.k = k; g.opa = OP_N; g.opb = OP_N;
                g.alpha = zddv{{1.0, 0.0}, {0.0, 0.0}}; g.beta = zddv{{0.5, 0.0}, {0.0, 0.0}};
                g.A = (const zddv *)A; g.lda = m; g.B = (const zddv *)B; g.ldb = k;
                g.C = (zddv *)C; g.ldc = m; g.variant = v;
                r2 = gemm_cplx_device(g, path, s);
            }
            float ms = 0.f;
            if (r2 || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = r2 ? r2 : -1;
                break;
            }
            if (rep > 0 && (t_v < 0 || ms < t_v)) t_v = ms;   // rep 0 warms up
        }
        if (!rc && all_ms && v < all_cap) all_ms[v] = t_v;
        if (!rc && (best < 0 || t_v < best_t)) { best = v; best_t = t_v; }
    }
    if (!rc && best >= 0) {
        if (path == MWGEMM_PATH_PLAIN) {
            tune_record(TK_PLAIN_REAL, m, n, k, best, best_t);
        } else {
            const int64_t mp = (m + 255) / 256 * 256, np = (n + 255) / 256 * 256;
            const int64_t kp = cplx ? 2 * ((k + 63) / 64 * 64) : (k + 63) / 64 * 64;
            tune_record(TK_OZ_GEMM, mp, np, kp, best, best_t);
        }
        if (variant) *variant = best;
        if (best_ms) *best_ms = best_t;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (C) (void)hipFree(C);
    return rc;
}

int mwgemm_batched_autotune(int64_t count, char transa, char transb, const int64_t *m,
                            const int64_t *n, const int64_t *k, int *choice, double *best_ms)
{
    if (count <= 0) return record_error(-3, "batched autotune: empty batch");
    const bool ta = toupper((unsigned char)transa) != 'N', tb = toupper((unsigned char)transb) != 'N';
    std::vector<int64_t> lda(count), ldb(count), ldc(count);
    std::vector<size_t> oa(count), ob(count), oc(count);
    size_t total = 0;
    double flops = 0.0;
    int64_t kmax = 0;
    for (int64_t p = 0; p < count; ++p) {
        const int64_t mm = std::max<int64_t>(m[p], 1), nn = std::max<int64_t>(n[p], 1), kk = std::max<int64_t>(k[p], 1);
        lda[p] = ta ? kk : mm;
        ldb[p] = tb ? nn : kk;
        ldc[p] = mm;
        oa[p] = total; total += (size_t)mm * kk;
        ob[p] = total; total += (size_t)kk * nn;
        oc[p] = total; total += (size_t)mm * nn;
        flops += 2.0 * (double)m[p] * (double)n[p] * (double)k[p];
        kmax = std::max(kmax, k[p]);
    }
    ddv *buf = nullptr;
    hipStream_t s = nullptr;
    int rc = 0;
    if (hipMalloc(&buf, total * sizeof(ddv)) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    {
        (void)hipGetLastError();
        rc = record_error(-6, "batched autotune: cannot allocate %zu bytes", total * sizeof(ddv));
    }
    if (!rc) rc = fill(buf, (int64_t)total, s);
    std::vector<const mw_float64x2 *> A(count), B(count);
    std::vector<mw_float64x2 *> C(count);
    std::vector<mw_float64x2> alpha(count, mw_float64x2{{1.0, 0.0}}), beta(count, mw_float64x2{{1.0, 0.0}});
    for (int64_t p = 0; p < count && buf; ++p) {
        A[p] = reinterpret_cast<const mw_float64x2 *>(buf + oa[p]);
        B[p] = reinterpret_cast<const mw_float64x2 *>(buf + ob[p]);
        C[p] = reinterpret_cast<mw_float64x2 *>(buf + oc[p]);
    }
    // candidates: plain; Ozaki with 128 / 64 residue tiles, or 64 with the reconstruction
    // fused (when the depth allows it)
    std::vector<int> cand = {0};
    if (kmax > 0 && ozaki_supported(kmax))
        for (int v = 0; v < BATCH_OZ_VARIANTS; ++v) cand.push_back(16 + v);
    int best = -1;
    double best_t = 0.0;
    for (int c : cand) {
        if (rc) break;
        const double t = batch_time_variant(count, transa, transb, m, n, k, alpha.data(), A.data(),
                                            lda.data(), B.data(), ldb.data(), beta.data(), C.data(),
                                            ldc.data(), c / 16 ? MWGEMM_PATH_OZAKI : MWGEMM_PATH_PLAIN,
                                            c % 16, 3, s);
        if (t < 0) { rc = (int)t; break; }
        if (best < 0 || t < best_t) { best = c; best_t = t; }
    }
    if (!rc && best >= 0) {
        const int64_t vol = (int64_t)std::cbrt(flops / 2.0 / (double)count);
        tune_record(TK_BATCH, count, vol, kmax, best, best_t);
        if (choice) *choice = best;
        if (best_ms) *best_ms = best_t;
    }
    if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
    if (buf) (void)hipFree(buf);
    return rc;
}

int mwgemm_autotune_lookup(int kind, int64_t a, int64_t b, int64_t c)
{
    if (kind < 0 || kind >= TK_COUNT) return -1;
    return tune_lookup(kind, a, b, c);
}

int mwgemm_autotune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    g_loaded = true;     // do not re-read the file after an explicit clear
    return 0;
}

}  // extern "C"
