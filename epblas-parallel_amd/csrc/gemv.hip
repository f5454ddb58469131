// gemv.hip -- double-double matrix-vector products y := alpha*op(A)*x + beta*y (mgemv_ /
// wgemv_), bit-for-bit the reference's AVX2 numerics (src/epblas-parallel/multifloats/
// mgemv.cpp:60-275, wgemv.cpp:55-300).  HBM-bound: every entry of A is read once
// (16 B real, 32 B complex, per multiply-add).
//
// op 'N': one lane per output row, the j loop sequential per row as in the reference (its
//   order decides the bits).  A workgroup (one wave) stages t_j = alpha*x_j (faithful),
//   the x_j == 0 skip flags and a per-batch "any skip" flag for 1024 columns at a time in
//   LDS; A is read column by column with consecutive lanes on consecutive rows.  With m
//   rows there are only m/64 waves, so once m is below ~64 rows per SIMD op 'N' is bound
//   by the latency of each row's chain (7 dependent FP64 adds per column step through the
//   sum's high limb), not by HBM.  Rows in whole
//   4-blocks of their band take the fast operations, the band tails the faithful ones (the
//   band layout is the reference's OpenMP row split for mgemv, and "whole 4-blocks, last
//   m % 4 rows" for wgemv).
// op 'T'/'C': four lanes per output column, lane L accumulating the rows i = L (mod 4)
//   below m - m%4 with the fast operations, reduced across the four lanes in the
//   reference's order (s0+s2, s1+s3, then their sum); lane 0 adds the last m % 4 rows and
//   alpha*s faithfully.
// In the scalar order (the reference's kernels for hosts without AVX2/FMA) every op-'N' row
// takes the faithful operations, and op 'T'/'C' is one sequential faithful dot per column.
// beta is applied to y in the same kernel, before the accumulation (mf_kernels.h:63-74).
#include "mw_internal.hpp"
#include "timer.hpp"

namespace mw {

namespace {

// Each lane's chain over A is sequential, so A entries are loaded in batches, the next
// batch in flight while the current one is consumed (software pipelining): a row (op 'N')
// or a quarter column (op 'T') per lane.  Batches are halved for complex entries, which
// take twice the registers.
template <typename V> constexpr int pf_n() { return sizeof(V) == sizeof(ddv) ? 16 : 8; }
template <typename V> constexpr int pf_t() { return sizeof(V) == sizeof(ddv) ? 8 : 4; }
constexpr int GEMV_CHUNK = 1024;   // columns whose alpha*x_j a workgroup stages at a time
// op 'N' workgroups are one wave: the rows (the only parallel dimension) then spread over
// every CU even for a few thousand rows.
constexpr int GEMV_NB = 64;

DDI int64_t vec_base(int64_t len, int64_t inc) { return inc < 0 ? -(len - 1) * inc : 0; }

// Whether row i is in the faithful tail of its band: bands follow blas_part_bound
// (common/blas_omp.h:79-84) over nbands parts; within a band the rows after the last whole
// 4-block are the tail.
DDI bool band_tail(int64_t i, int64_t m, int64_t nbands)
{
    const int64_t base = m / nbands, rem = m % nbands;
    int64_t t;
    if (i < rem * (base + 1)) t = i / (base + 1);
    else t = rem + (i - rem * (base + 1)) / base;
    const int64_t lo = t * base + (t < rem ? t : rem);
    const int64_t len = base + (t < rem ? 1 : 0);
    return i - lo >= len / 4 * 4;
}

__host__ __device__ inline bool is0(const ddv &v) { return v.hi == 0.0 && v.lo == 0.0; }
__host__ __device__ inline bool is0(const zddv &v) { return is0(v.re) && is0(v.im); }
DDI ddv prepass(const ddv &y, const ddv &b, int mode) { return dd_beta_prepass(y, b, mode); }
DDI zddv prepass(const zddv &y, const zddv &b, int mode) { return zdd_beta_prepass(y, b, mode); }
DDI ddv mul_exact(const ddv &a, const ddv &b) { return dd_mul_exact(a, b); }
DDI zddv mul_exact(const zddv &a, const zddv &b) { return zdd_mul_exact(a, b); }
DDI ddv add_exact(const ddv &a, const ddv &b) { return dd_add_exact(a, b); }
DDI zddv add_exact(const zddv &a, const zddv &b) { return zdd_add_exact(a, b); }
DDI ddv mul_fast(const ddv &a, const ddv &b) { return dd_mul_fast(a, b); }
DDI zddv mul_fast(const zddv &a, const zddv &b) { return zdd_mul_fast(a, b); }
DDI ddv add_fast(const ddv &a, const ddv &b) { return dd_add_fast(a, b); }
DDI zddv add_fast(const zddv &a, const zddv &b) { return zdd_add_fast(a, b); }
DDI ddv conj_if(const ddv &a, bool) { return a; }
DDI zddv conj_if(const zddv &a, bool c) { return c ? zdd_conj(a) : a; }
DDI ddv zero_of(ddv) { return ddv{0.0, 0.0}; }
DDI zddv zero_of(zddv) { return zddv{{0.0, 0.0}, {0.0, 0.0}}; }

DDI ddv shfl_xor(const ddv &v, int mask)
{
    return ddv{__shfl_xor(v.hi, mask, 64), __shfl_xor(v.lo, mask, 64)};
}
DDI zddv shfl_xor(const zddv &v, int mask) { return zddv{shfl_xor(v.re, mask), shfl_xor(v.im, mask)}; }

// One step of a row's chain: y := y + t*a (fast operations, or faithful for tail rows).
// With SKIP, a column with x_j == 0 keeps y (a NaN or Inf in such a column must not reach
// y), as a select so that the batch stays one straight-line block.
template <bool TAIL, bool SKIP, typename V>
DDI V row_step(const V &yi, const V &t, const V &a, int skip)
{
    const V upd = TAIL ? add_exact(yi, mul_exact(t, a)) : add_fast(yi, mul_fast(t, a));
    return (SKIP && skip) ? yi : upd;
}

// The steps of one batch of pf_n columns whose A entries are in registers; the selects
// are paid only in a batch that holds a zero x_j (the flag is the same for the whole wave).
template <bool TAIL, typename V>
DDI V row_batch(V yi, const V *av, const V *ts, const int *skip, int any_skip)
{
    constexpr int PF = pf_n<V>();
    if (__builtin_amdgcn_readfirstlane(any_skip)) {
#pragma unroll
        for (int q = 0; q < PF; ++q) yi = row_step<TAIL, true>(yi, ts[q], av[q], skip[q]);
    } else {
#pragma unroll
        for (int q = 0; q < PF; ++q) yi = row_step<TAIL, false>(yi, ts[q], av[q], 0);
    }
    return yi;
}

// Loads the A entries of batch b of the chunk: the column base is the same for the whole
// wave, the row offset i is the lane's.
template <typename V>
DDI void row_load(V *av, const V *__restrict__ acol, int64_t lda, int b, int64_t i)
{
    constexpr int PF = pf_n<V>();
    const V *base = acol + (int64_t)b * PF * lda;
#pragma unroll
    for (int q = 0; q < PF; ++q) av[q] = (base + (int64_t)q * lda)[i];
}

// The column chunk [j0, j0 + cnt) of row i: full batches of pf_n loads, each batch's loads
// issued before the previous batch's steps (two register sets used in turn, no branches
// inside a batch), then the remainder one by one.
template <bool TAIL, typename V>
DDI V row_chunk(V yi, const V *__restrict__ acol, int64_t lda, int64_t i, int cnt,
                const V *ts, const int *skip, const int *bskip)
{
    constexpr int PF = pf_n<V>();
    const int nb = cnt / PF;
    if (nb > 0) {
        V ba[PF], bb[PF];
        row_load(ba, acol, lda, 0, i);
        int b = 0;
        for (; b + 2 < nb; b += 2) {
            row_load(bb, acol, lda, b + 1, i);
            __builtin_amdgcn_sched_barrier(0);   // next batch in flight before these steps
            yi = row_batch<TAIL>(yi, ba, ts + b * PF, skip + b * PF, bskip[b]);
            row_load(ba, acol, lda, b + 2, i);
            __builtin_amdgcn_sched_barrier(0);
            yi = row_batch<TAIL>(yi, bb, ts + (b + 1) * PF, skip + (b + 1) * PF, bskip[b + 1]);
        }
        if (b + 1 < nb) {
            row_load(bb, acol, lda, b + 1, i);
            __builtin_amdgcn_sched_barrier(0);
            yi = row_batch<TAIL>(yi, ba, ts + b * PF, skip + b * PF, bskip[b]);
            yi = row_batch<TAIL>(yi, bb, ts + (b + 1) * PF, skip + (b + 1) * PF, bskip[b + 1]);
        } else {
            yi = row_batch<TAIL>(yi, ba, ts + b * PF, skip + b * PF, bskip[b]);
        }
    }
    for (int q = nb * PF; q < cnt; ++q)
        yi = row_step<TAIL, true>(yi, ts[q], (acol + (int64_t)q * lda)[i], skip[q]);
    return yi;
}

// op 'N'.  nbands <= 0: the complex layout (only the last m % 4 rows are tail rows).
template <typename V>
__global__ __launch_bounds__(GEMV_NB)
void k_gemv_n(int64_t m, int64_t n, V alpha, V beta, int beta_mode, int alpha0,
              const V *__restrict__ A, int64_t lda, const V *__restrict__ x, int64_t incx,
              V *__restrict__ y, int64_t incy, int64_t nbands, int all_tail)
{
    constexpr int PF = pf_n<V>();
    __shared__ V ts[GEMV_CHUNK];
    __shared__ int skip[GEMV_CHUNK];
    __shared__ int bskip[GEMV_CHUNK / PF];   // per batch: does it hold a zero x_j
    const int64_t i = (int64_t)blockIdx.x * GEMV_NB + threadIdx.x;
    const bool live = i < m;
    const int64_t by = vec_base(m, incy), bx = vec_base(n, incx);
    // all_tail: the scalar order, every row takes the faithful operations
    const bool tail = live && (all_tail || (nbands > 0 ? band_tail(i, m, nbands) : i >= m / 4 * 4));
    V yi = zero_of(alpha);
    if (live) yi = prepass(y[by + i * incy], beta, beta_mode);
    if (!alpha0) {
        for (int64_t j0 = 0; j0 < n; j0 += GEMV_CHUNK) {
            __syncthreads();
            for (int c = threadIdx.x; c < GEMV_CHUNK; c += GEMV_NB) {
                const int64_t jj = j0 + c;
                if (jj < n) {
                    const V xj = x[bx + jj * incx];
                    skip[c] = is0(xj) ? 1 : 0;
                    ts[c] = mul_exact(alpha, xj);
                }
            }
            __syncthreads();
            const int cnt = (int)(n - j0 < GEMV_CHUNK ? n - j0 : GEMV_CHUNK);
            for (int b = threadIdx.x; b < GEMV_CHUNK / PF; b += GEMV_NB) {
                int f = 0;
                for (int q = b * PF; q < b * PF + PF && q < cnt; ++q) f |= skip[q];
                bskip[b] = f;
            }
            __syncthreads();
            if (!live) continue;
            const V *acol = A + j0 * lda;
            yi = tail ? row_chunk<true>(yi, acol, lda, i, cnt, ts, skip, bskip)
                      : row_chunk<false>(yi, acol, lda, i, cnt, ts, skip, bskip);
        }
    }
    if (live) y[by + i * incy] = yi;
}

// op 'T' / 'C': 64 columns per workgroup, four lanes each.
template <typename V, bool CONJ>
__global__ __launch_bounds__(256)
void k_gemv_t(int64_t m, int64_t n, V alpha, V beta, int beta_mode, int alpha0,
              const V *__restrict__ A, int64_t lda, const V *__restrict__ x, int64_t incx,
              V *__restrict__ y, int64_t incy)
{
    const int64_t j = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2);
    const int L = threadIdx.x & 3;
    const bool live = j < n;
    const int64_t m4 = m / 4 * 4;
    const int64_t bx = vec_base(m, incx), by = vec_base(n, incy);
    const V *aj = A + (live ? j : 0) * lda;
    V s = zero_of(alpha);
    if (live && !alpha0) {
        constexpr int PF = pf_t<V>();
        int64_t i0 = L;
        // batches of PF rows of this lane (i0, i0 + 4, ...), the next batch in flight
        if (i0 + 4 * (PF - 1) < m4) {
            V ca[PF], cx[PF], na[PF], nx[PF];
#pragma unroll
            for (int q = 0; q < PF; ++q) {
                ca[q] = aj[i0 + 4 * q];
                cx[q] = x[bx + (i0 + 4 * q) * incx];
            }
            for (; i0 + 4 * PF + 4 * (PF - 1) < m4; i0 += 4 * PF) {
#pragma unroll
                for (int q = 0; q < PF; ++q) {
                    na[q] = aj[i0 + 4 * PF + 4 * q];
                    nx[q] = x[bx + (i0 + 4 * PF + 4 * q) * incx];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < PF; ++q) s = add_fast(s, mul_fast(conj_if(ca[q], CONJ), cx[q]));
#pragma unroll
                for (int q = 0; q < PF; ++q) {
                    ca[q] = na[q];
                    cx[q] = nx[q];
                }
            }
#pragma unroll
            for (int q = 0; q < PF; ++q) s = add_fast(s, mul_fast(conj_if(ca[q], CONJ), cx[q]));
            i0 += 4 * PF;
        }
        for (; i0 < m4; i0 += 4) s = add_fast(s, mul_fast(conj_if(aj[i0], CONJ), x[bx + i0 * incx]));
    }
    // reference order: (s0 + s2, s1 + s3), then lane 0 + lane 1
    const V p = add_fast(s, shfl_xor(s, 2));
    V r = add_fast(p, shfl_xor(p, 1));
    if (live && L == 0) {
        V yj = prepass(y[by + j * incy], beta, beta_mode);
        if (!alpha0) {
            for (int64_t i = m4; i < m; ++i) r = add_exact(r, mul_exact(conj_if(aj[i], CONJ), x[bx + i * incx]));
            yj = add_exact(yj, mul_exact(alpha, r));
        }
        y[by + j * incy] = yj;
    }
}

// op 'T' / 'C' in the scalar order (the reference's kernels for hosts without AVX2/FMA):
// one sequential faithful chain per column, s := s + a_ij * x_i for i ascending, then
// y_j := y_j + alpha * s (mgemv.cpp:213-223, wgemv.cpp:284-295).  A one-wave workgroup owns
// 64 columns; blocks of ROWS rows of them are staged in LDS by coalesced column reads
// (consecutive lanes on consecutive rows), then each lane walks its own column.
template <typename V, bool CONJ>
__global__ __launch_bounds__(64)
void k_gemv_t_seq(int64_t m, int64_t n, V alpha, V beta, int beta_mode, int alpha0,
                  const V *__restrict__ A, int64_t lda, const V *__restrict__ x, int64_t incx,
                  V *__restrict__ y, int64_t incy)
{
    constexpr int ROWS = sizeof(V) == sizeof(ddv) ? 32 : 16;
    __shared__ V tile[64][ROWS + 1];   // [column][row]; the odd pitch spreads the banks
    __shared__ V xs[ROWS];
    const int lane = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.x * 64, j = j0 + lane;
    const bool live = j < n;
    const int64_t bx = vec_base(m, incx), by = vec_base(n, incy);
    V s = zero_of(alpha);
    if (!alpha0) {
        const int cols = (int)(n - j0 < 64 ? n - j0 : 64);
        for (int64_t r0 = 0; r0 < m; r0 += ROWS) {
            const int rows = (int)(m - r0 < ROWS ? m - r0 : ROWS);
            __syncthreads();   // the previous block has been consumed
            if (lane < rows) {
                for (int c = 0; c < cols; ++c) tile[c][lane] = A[(j0 + c) * lda + r0 + lane];
                xs[lane] = x[bx + (r0 + lane) * incx];
            }
            __syncthreads();
            if (live)
                for (int r = 0; r < rows; ++r)
                    s = add_exact(s, mul_exact(conj_if(tile[lane][r], CONJ), xs[r]));
        }
    }
    if (live) {
        V yj = prepass(y[by + j * incy], beta, beta_mode);
        if (!alpha0) yj = add_exact(yj, mul_exact(alpha, s));
        y[by + j * incy] = yj;
    }
}

template <typename V>
hipError_t launch_gemv(int op, int64_t m, int64_t n, V alpha, V beta, int beta_mode, bool alpha0,
                       const V *A, int64_t lda, const V *x, int64_t incx, V *y, int64_t incy,
                       int64_t nbands, int order, hipStream_t s)
{
    ScopedTimer tm(KC_GEMV, s);
    const int a0 = alpha0 ? 1 : 0;
    if (op == OP_N) {
        hipLaunchKernelGGL((k_gemv_n<V>), dim3((unsigned)((m + GEMV_NB - 1) / GEMV_NB)), dim3(GEMV_NB), 0, s, m, n,
                           alpha, beta, beta_mode, a0, A, lda, x, incx, y, incy, nbands,
                           order == ORDER_SCALAR ? 1 : 0);
    } else if (order == ORDER_SCALAR) {
        const dim3 g((unsigned)((n + 63) / 64));
        if (op == OP_C)
            hipLaunchKernelGGL((k_gemv_t_seq<V, true>), g, dim3(64), 0, s, m, n, alpha, beta,
                               beta_mode, a0, A, lda, x, incx, y, incy);
        else
            hipLaunchKernelGGL((k_gemv_t_seq<V, false>), g, dim3(64), 0, s, m, n, alpha, beta,
                               beta_mode, a0, A, lda, x, incx, y, incy);
    } else if (op == OP_C) {
        hipLaunchKernelGGL((k_gemv_t<V, true>), dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, m,
                           n, alpha, beta, beta_mode, alpha0 ? 1 : 0, A, lda, x, incx, y, incy);
    } else {
        hipLaunchKernelGGL((k_gemv_t<V, false>), dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s,
                           m, n, alpha, beta, beta_mode, alpha0 ? 1 : 0, A, lda, x, incx, y, incy);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t gemv_real(int op, int64_t m, int64_t n, ddv alpha, ddv beta, const ddv *A, int64_t lda,
                     const ddv *x, int64_t incx, ddv *y, int64_t incy, int64_t nbands,
                     int order, hipStream_t s)
{
    if (m == 0 || n == 0) return hipSuccess;
    const int mode = is0(beta) ? BETA_ZERO : (beta.hi == 1.0 && beta.lo == 0.0) ? BETA_ONE : BETA_SCALE;
    return launch_gemv<ddv>(op == OP_C ? OP_T : op, m, n, alpha, beta, mode, is0(alpha), A, lda,
                            x, incx, y, incy, nbands < 1 ? 1 : nbands, order, s);
}

hipError_t gemv_cplx(int op, int64_t m, int64_t n, zddv alpha, zddv beta, const zddv *A,
                     int64_t lda, const zddv *x, int64_t incx, zddv *y, int64_t incy, int order,
                     hipStream_t s)
{
    if (m == 0 || n == 0) return hipSuccess;
    const bool b1 = beta.re.hi == 1.0 && beta.re.lo == 0.0 && is0(beta.im);
    const int mode = is0(beta) ? BETA_ZERO : b1 ? BETA_ONE : BETA_SCALE;
    return launch_gemv<zddv>(op, m, n, alpha, beta, mode, is0(alpha), A, lda, x, incx, y, incy,
                             0, order, s);
}

}  // namespace mw
