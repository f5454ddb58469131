// plain_tile.hpp -- device-side building blocks of the plain double-double GEMM: slab
// staging (global -> FP64-vector registers -> LDS, transposition and conjugation folded in)
// and the per-tile K loop, shared by the single-GEMM kernels and the persistent batched one.
#pragma once
#include "mw_internal.hpp"

namespace mw {

constexpr int KCHUNK = 128;   // reference KC: the accumulator flush period

// ----------------------------------------------------------------------------------
// Tile staging helpers (generic over element type T = ddv or zddv).
// ----------------------------------------------------------------------------------

// Staging registers hold native FP64 vectors (a DD as double2, a complex DD as double4)
// so they stay in VGPRs; arrays of the element structs end up in scratch memory.
typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
template <typename T> struct VecOf;
template <> struct VecOf<ddv> { typedef v2d type; };
template <> struct VecOf<zddv> { typedef v4d type; };
DDI v2d conj_vec(v2d x) { return x; }
DDI v4d conj_vec(v4d x) { return v4d{x.x, x.y, -x.z, -x.w}; }

// op(A) slab [i0, i0+BM) x [k0, k0+BK) -> regs.  TRANS: A is stored K x M.
template <typename T, int BM, int BK, bool TRANS, bool CONJ>
struct SlabA {
    typedef typename VecOf<T>::type V;
    static constexpr int PER = BM * BK / 256;
    V v[PER];
    DDI void load(const T *__restrict__ A, int64_t lda, int64_t m, int64_t k, int64_t i0,
                  int64_t k0, int tid)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + 256 * q;
            int r, c;
            if (!TRANS) { r = e % BM; c = e / BM; }
            else        { c = e % BK; r = e / BK; }
            int64_t gi = i0 + r, gk = k0 + c;
            V x = V(0.0);
            if (gi < m && gk < k) {
                x = *reinterpret_cast<const V *>(TRANS ? &A[gi * lda + gk] : &A[gk * lda + gi]);
                if constexpr (CONJ) x = conj_vec(x);
            }
            v[q] = x;
        }
    }
    template <int LDW>
    DDI void store(T (*S)[LDW], int tid) const
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + 256 * q;
            int r, c;
            if (!TRANS) { r = e % BM; c = e / BM; }
            else        { c = e % BK; r = e / BK; }
            *reinterpret_cast<V *>(&S[c][r]) = v[q];
        }
    }
};

// op(B) slab [k0, k0+BK) x [j0, j0+BN) -> regs.  TRANS: B is stored N x K.
template <typename T, int BN, int BK, bool TRANS, bool CONJ>
struct SlabB {
    typedef typename VecOf<T>::type V;
    static constexpr int PER = BN * BK / 256;
    V v[PER];
    DDI void load(const T *__restrict__ B, int64_t ldb, int64_t n, int64_t k, int64_t j0,
                  int64_t k0, int tid)
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + 256 * q;
            int r, c;   // r: column offset in [0,BN), c: k offset in [0,BK)
            if (!TRANS) { c = e % BK; r = e / BK; }
            else        { r = e % BN; c = e / BN; }
            int64_t gj = j0 + r, gk = k0 + c;
            V x = V(0.0);
            if (gj < n && gk < k) {
                x = *reinterpret_cast<const V *>(TRANS ? &B[gk * ldb + gj] : &B[gj * ldb + gk]);
                if constexpr (CONJ) x = conj_vec(x);
            }
            v[q] = x;
        }
    }
    template <int LDW>
    DDI void store(T (*S)[LDW], int tid) const
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + 256 * q;
            int r, c;
            if (!TRANS) { c = e % BK; r = e / BK; }
            else        { r = e % BN; c = e / BN; }
            *reinterpret_cast<V *>(&S[c][r]) = v[q];
        }
    }
};

// One BM x BN tile of a real plain GEMM (the whole K loop with the reference's
// 128-deep chunk flushes); As/Bs are the workgroup's LDS slabs.  Ends with a barrier, so the
// caller may reuse the LDS for another tile right away.
template <int TM, int TN, int BK, bool TA, bool TB>
DDI void mgemm_plain_tile(int64_t m, int64_t n, int64_t k, ddv alpha, ddv beta, int beta_mode,
                          const ddv *__restrict__ A, int64_t lda, const ddv *__restrict__ B,
                          int64_t ldb, ddv *__restrict__ C, int64_t ldc, int64_t i0, int64_t j0,
                          ddv (&As)[BK][16 * TM], ddv (&Bs)[BK][16 * TN])
{
    constexpr int BM = 16 * TM, BN = 16 * TN;
    static_assert(KCHUNK % BK == 0, "BK must divide the flush period");
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    SlabA<ddv, BM, BK, TA, false> sa;
    SlabB<ddv, BN, BK, TB, false> sb;

    ddv acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = ddv{0.0, 0.0};

    const int64_t ntiles = (k + BK - 1) / BK;
    sa.load(A, lda, m, k, i0, 0, tid);
    sb.load(B, ldb, n, k, j0, 0, tid);
    sa.store(As, tid);
    sb.store(Bs, tid);
    __syncthreads();

    for (int64_t t = 0; t < ntiles; ++t) {
        const int64_t k0 = t * BK;
        if (t + 1 < ntiles) {
            sa.load(A, lda, m, k, i0, k0 + BK, tid);
            sb.load(B, ldb, n, k, j0, k0 + BK, tid);
        }
        const int kmax = (k - k0 < BK) ? (int)(k - k0) : BK;
        if (kmax == BK) {
#pragma unroll 2
            for (int kk = 0; kk < BK; ++kk) {
                ddv a[TM], b[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
                for (int j = 0; j < TN; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = dd_add_fast(acc[i][j], dd_mul_fast(a[i], b[j]));
            }
        } else {
            for (int kk = 0; kk < kmax; ++kk) {
                ddv a[TM], b[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
                for (int j = 0; j < TN; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = dd_add_fast(acc[i][j], dd_mul_fast(a[i], b[j]));
            }
        }
        const int64_t kend = k0 + kmax;
        if (kend % KCHUNK == 0 || kend == k) {
            const bool first = (kend <= KCHUNK);
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int64_t gi = i0 + tx + 16 * i;
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const int64_t gj = j0 + ty + 16 * j;
                    if (gi < m && gj < n) {
                        ddv *cp = &C[gj * ldc + gi];
                        ddv c = *cp;
                        if (first) c = dd_beta_prepass(c, beta, beta_mode);
                        *cp = dd_add_exact(c, dd_mul_fast(alpha, acc[i][j]));
                    }
                    acc[i][j] = ddv{0.0, 0.0};
                }
            }
        }
        __syncthreads();
        if (t + 1 < ntiles) {
            sa.store(As, tid);
            sb.store(Bs, tid);
            __syncthreads();
        }
    }
}

}  // namespace mw
