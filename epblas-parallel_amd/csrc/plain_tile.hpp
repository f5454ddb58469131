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
DDI ddv mul_exact_t(const ddv &a, const ddv &b) { return dd_mul_exact(a, b); }
DDI zddv mul_exact_t(const zddv &a, const zddv &b) { return zdd_mul_exact(a, b); }
DDI ddv add_exact_t(const ddv &a, const ddv &b) { return dd_add_exact(a, b); }
DDI zddv add_exact_t(const zddv &a, const zddv &b) { return zdd_add_exact(a, b); }
DDI ddv beta_prepass_t(const ddv &c, const ddv &beta, int mode) { return dd_beta_prepass(c, beta, mode); }
DDI zddv beta_prepass_t(const zddv &c, const zddv &beta, int mode) { return zdd_beta_prepass(c, beta, mode); }

// Plain-path operation orders.  ORDER_SIMD: the reference's AVX2 kernel (fast operations
// accumulated per 128-deep chunk, then alpha and a faithful add into C; mgemm_serial.cpp:
// 173-319).  ORDER_SCALAR: its scalar kernel for hosts without AVX2/FMA (every
// multiply-add straight into C with the faithful operations, C += (alpha*b)*a;
// mgemm_serial.cpp:87-99, wgemm_serial.cpp:102-115).

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
    // As store, each entry replaced by alpha (x) entry with the faithful DD (or complex DD)
    // multiply: the reference's scalar-order kernel forms t = alpha * b per B entry.
    template <int LDW>
    DDI void store_scaled(T (*S)[LDW], const T &alpha, int tid) const
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int e = tid + 256 * q;
            int r, c;
            if (!TRANS) { c = e % BK; r = e / BK; }
            else        { r = e % BN; c = e / BN; }
            T x;
            __builtin_memcpy(&x, &v[q], sizeof(T));
            S[c][r] = mul_exact_t(alpha, x);
        }
    }
};

// Triangular-output GEMM (gemmtr) masks.  The n x n C is walked in column blocks of nb;
// for column j in block [jc, jc + nb) the UPLO triangle splits into the block's diagonal
// triangle (rows jc..j for 'U', j..jc+nb-1 for 'L', updated by the scalar kernel) and the
// rectangle beyond it (rows i < jc for 'U', i >= jc + nb for 'L'), which takes the GEMM's
// own per-element sequence (mgemmtr_serial.cpp:117-166).  TRI_NONE is the plain GEMM.
enum { TRI_NONE = 0, TRI_UPPER = 1, TRI_LOWER = 2 };

template <int TRI> DDI bool tri_rect(int64_t i, int64_t j, int64_t nb)
{
    if constexpr (TRI == TRI_UPPER) return i < (j / nb) * nb;
    else if constexpr (TRI == TRI_LOWER) return i >= (j / nb) * nb + nb;
    else return true;
}

// Whether the BM x BN tile at (i0, j0) of an n x n C holds a rectangle element.
template <int TRI> DDI bool tri_tile_live(int64_t i0, int64_t j0, int BM, int BN, int64_t n, int64_t nb)
{
    if constexpr (TRI == TRI_UPPER) {
        const int64_t jl = (j0 + BN < n ? j0 + BN : n) - 1;
        return i0 < (jl / nb) * nb;
    } else if constexpr (TRI == TRI_LOWER) {
        const int64_t il = (i0 + BM < n ? i0 + BM : n) - 1;
        return il >= (j0 / nb) * nb + nb;
    } else {
        return true;
    }
}

// Live rectangle tiles of tile column jt (BM x BN tiles over the n x n C, tiles_m tile rows):
// tile rows first .. first + count - 1, exactly those for which tri_tile_live holds.
//   'U': live while i0 < R, R = ((min(j0 + BN, n) - 1) / nb) * nb, so rows [0, ceil(R / BM)).
//   'L': live once i0 + BM - 1 (or n - 1 for the last row) >= X = (j0 / nb) * nb + nb, so
//        rows [X / BM, tiles_m) when X <= n - 1 and none otherwise.
// 32-bit arithmetic: n < 2^31, and nb is clamped to n + 1 (the same masks as any larger nb).
template <int TRI>
__host__ __device__ inline void tri_live_rows(int jt, int BM, int BN, int n, int nb, int tiles_m,
                                              int &first, int &count)
{
    const int j0 = jt * BN;
    if (TRI == TRI_UPPER) {
        const int jl = (j0 + BN < n ? j0 + BN : n) - 1;
        const int R = (jl / nb) * nb;
        const int c = (R + BM - 1) / BM;
        first = 0;
        count = c < tiles_m ? c : tiles_m;
    } else {
        const int X = (j0 / nb) * nb + nb;
        first = X <= n - 1 ? X / BM : 0;
        count = X <= n - 1 ? tiles_m - X / BM : 0;
    }
}

// The idx-th live rectangle tile, counted column by column (tile columns in ascending order,
// live rows ascending within a column): the gemmtr grid has one workgroup per live tile
// instead of one per tile of the n x n square.  Each wave scans 64 tile columns per step
// (one per lane: live counts, inclusive prefix sum, then the first lane whose prefix passes
// idx); every wave of the workgroup computes the same (i0, j0).
template <int TRI, int BM, int BN>
DDI void tri_live_tile(int idx, int n, int nb, int tiles_m, int64_t &i0, int64_t &j0)
{
    const int lane = threadIdx.x & 63;
    const int tn = (n + BN - 1) / BN;
    int before = 0;                      // live tiles in the columns of earlier steps
    for (int c0 = 0; c0 < tn; c0 += 64) {
        const int c = c0 + lane;
        int first = 0, cnt = 0;
        if (c < tn) tri_live_rows<TRI>(c, BM, BN, n, nb, tiles_m, first, cnt);
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int target = idx - before;
        const unsigned long long hit = __ballot(c < tn && target < incl);
        if (hit) {
            const int L = __ffsll((long long)hit) - 1;
            const int f = __shfl(first, L, 64), ex = __shfl(incl - cnt, L, 64);
            i0 = (int64_t)(f + target - ex) * BM;
            j0 = (int64_t)(c0 + L) * BN;
            return;
        }
        before += __shfl(incl, 63, 64);
    }
    i0 = (int64_t)tiles_m * BM;           // idx beyond the live tiles (not launched)
    j0 = 0;
}

// The lane's TM x TN register-tile elements that a tile writes: in range and, for gemmtr,
// in the rectangle; fixed for the tile, so one bit mask (one register) for the K loop.
template <int TM, int TN, int TRI>
DDI uint32_t tile_keep(int64_t m, int64_t n, int64_t i0, int64_t j0, int tx, int ty, int64_t tri_nb)
{
    static_assert(TM * TN <= 32, "store mask width");
    uint32_t keep = 0;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int64_t gi = i0 + tx + 16 * i, gj = j0 + ty + 16 * j;
            if (gi < m && gj < n && tri_rect<TRI>(gi, gj, tri_nb)) keep |= 1u << (i * TN + j);
        }
    return keep;
}

// One tile in the scalar order (real or complex elements T): C is pre-passed with beta
// and kept in registers, and for k ascending C := C + t_kj * a_ik with t_kj = alpha * b_kj
// formed once per B entry as the slab is staged (faithful operations throughout).
template <int TM, int TN, int BK, bool OA_T, bool OB_T, int TRI, bool OA_C = false,
          bool OB_C = false, typename T>
DDI void scalar_order_tile(int64_t m, int64_t n, int64_t k, T alpha, T beta, int beta_mode,
                           const T *__restrict__ A, int64_t lda, const T *__restrict__ B,
                           int64_t ldb, T *__restrict__ C, int64_t ldc, int64_t i0, int64_t j0,
                           T (&As)[BK][16 * TM], T (&Bs)[BK][16 * TN], int64_t tri_nb)
{
    constexpr int BM = 16 * TM, BN = 16 * TN;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    SlabA<T, BM, BK, OA_T, OA_C> sa;
    SlabB<T, BN, BK, OB_T, OB_C> sb;
    const uint32_t keep = tile_keep<TM, TN, TRI>(m, n, i0, j0, tx, ty, tri_nb);

    T c[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int64_t gi = i0 + tx + 16 * i, gj = j0 + ty + 16 * j;
            T v{};
            if ((keep >> (i * TN + j)) & 1u) v = beta_prepass_t(C[gj * ldc + gi], beta, beta_mode);
            c[i][j] = v;
        }

    const int64_t ntiles = (k + BK - 1) / BK;
    sa.load(A, lda, m, k, i0, 0, tid);
    sb.load(B, ldb, n, k, j0, 0, tid);
    sa.store(As, tid);
    sb.store_scaled(Bs, alpha, tid);
    __syncthreads();
    for (int64_t t = 0; t < ntiles; ++t) {
        const int64_t k0 = t * BK;
        if (t + 1 < ntiles) {
            sa.load(A, lda, m, k, i0, k0 + BK, tid);
            sb.load(B, ldb, n, k, j0, k0 + BK, tid);
        }
        const int kmax = (k - k0 < BK) ? (int)(k - k0) : BK;
        for (int kk = 0; kk < kmax; ++kk) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) c[i][j] = add_exact_t(c[i][j], mul_exact_t(b[j], a[i]));
        }
        __syncthreads();
        if (t + 1 < ntiles) {
            sa.store(As, tid);
            sb.store_scaled(Bs, alpha, tid);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
            if ((keep >> (i * TN + j)) & 1u) C[(j0 + ty + 16 * j) * ldc + i0 + tx + 16 * i] = c[i][j];
}

template <int TM, int TN, int BK, bool TA, bool TB, int TRI>
DDI void mgemm_scalar_tile(int64_t m, int64_t n, int64_t k, ddv alpha, ddv beta, int beta_mode,
                           const ddv *__restrict__ A, int64_t lda, const ddv *__restrict__ B,
                           int64_t ldb, ddv *__restrict__ C, int64_t ldc, int64_t i0, int64_t j0,
                           ddv (&As)[BK][16 * TM], ddv (&Bs)[BK][16 * TN], int64_t tri_nb)
{
    scalar_order_tile<TM, TN, BK, TA, TB, TRI>(m, n, k, alpha, beta, beta_mode, A, lda, B, ldb, C,
                                               ldc, i0, j0, As, Bs, tri_nb);
}

// One BM x BN tile of a real plain GEMM (the whole K loop, in the operation order SCALAR
// selects: 128-deep chunk flushes, or every multiply-add into C); As/Bs are the workgroup's
// LDS slabs.  Ends with a barrier, so the caller may reuse the LDS for another tile right
// away.  With TRI, only the gemmtr rectangle elements of the tile are read and written.
template <int TM, int TN, int BK, bool TA, bool TB, int TRI = TRI_NONE, bool SCALAR = false>
DDI void mgemm_plain_tile(int64_t m, int64_t n, int64_t k, ddv alpha, ddv beta, int beta_mode,
                          const ddv *__restrict__ A, int64_t lda, const ddv *__restrict__ B,
                          int64_t ldb, ddv *__restrict__ C, int64_t ldc, int64_t i0, int64_t j0,
                          ddv (&As)[BK][16 * TM], ddv (&Bs)[BK][16 * TN], int64_t tri_nb = 0)
{
    if constexpr (SCALAR) {
        mgemm_scalar_tile<TM, TN, BK, TA, TB, TRI>(m, n, k, alpha, beta, beta_mode, A, lda, B, ldb,
                                                   C, ldc, i0, j0, As, Bs, tri_nb);
        return;
    }
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

    const uint32_t keep = tile_keep<TM, TN, TRI>(m, n, i0, j0, tx, ty, tri_nb);

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
                    if ((keep >> (i * TN + j)) & 1u) {
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
