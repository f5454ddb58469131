// plain_gemm.hip -- the "plain" double-double GEMMs on the FP64 vector ALU.
//
// One workgroup (256 lanes = 4 wave64) owns a BM x BN tile of C; each lane owns a TM x TN
// register tile whose rows are strided by 16 (lane%16 + 16*i) and columns by 16
// (lane/16 + 16*j), so every LDS read is either conflict-free or a broadcast.  op(A) and
// op(B) K-slabs of depth BK are staged global -> registers -> LDS (transpose and
// conjugation folded into the staging), with the next slab's global loads in flight while
// the current slab is consumed.
//
// Per output element the arithmetic is exactly the reference's (bit-for-bit):
//   C := beta-prepass(C)                      (multifloats/mgemm_parallel.cpp:68-77)
//   for each K chunk of 128 from k = 0        (mgemm_serial.cpp:35-37, :126)
//       acc := (0,0); acc := fast_add(acc, fast_mul(a_ip, b_pj)), p ascending
//                                             (mgemm_serial.cpp:241-269)
//       C := faithful_add(C, fast_mul(alpha, acc))   (mgemm_serial.cpp:173-191)
// The running C lives in HBM between chunk flushes: one read-modify-write per 128-deep
// chunk costs ~1 byte per CU-clock, far below what the FP64 ALU leaves idle.
// The reference's scalar kernel (hosts without AVX2/FMA) is the second order, SCALAR:
//   C := beta-prepass(C); for k ascending: C := faithful_add(C, faithful_mul(alpha*b_kj, a_ik))
//                                             (mgemm_serial.cpp:87-99; wgemm_serial.cpp:102-115)
#include <algorithm>
#include <atomic>

#include "plain_tile.hpp"
#include "timer.hpp"

namespace mw {

// ----------------------------------------------------------------------------------
// Triangular-output GEMM (gemmtr): the nb x nb diagonal triangles, one lane per element,
// with the reference's scalar faithful sequence (mgemmtr_serial.cpp:48-95 / complex
// wgemmtr_serial.cpp:48-95): the beta scaling, then for op(A) = 'N'
//     per l ascending: t = alpha * b_lj; skipped when t == 0; c += t * a_il
// and otherwise  s = sum_l a_li * b_lj (ascending, faithful), c += alpha * s.
// Lane (j, r): column j, r-th row of its block's triangle part of that column.
// ----------------------------------------------------------------------------------
DDI ddv mul_exact_any(const ddv &a, const ddv &b) { return dd_mul_exact(a, b); }
DDI zddv mul_exact_any(const zddv &a, const zddv &b) { return zdd_mul_exact(a, b); }
DDI ddv add_exact_any(const ddv &a, const ddv &b) { return dd_add_exact(a, b); }
DDI zddv add_exact_any(const zddv &a, const zddv &b) { return zdd_add_exact(a, b); }
DDI bool is_zero_any(const ddv &v) { return v.hi == 0.0 && v.lo == 0.0; }
DDI bool is_zero_any(const zddv &v) { return is_zero_any(v.re) && is_zero_any(v.im); }
// p ? a : b field by field (a select of whole structs goes through scratch memory).
DDI ddv select_any(bool p, const ddv &a, const ddv &b)
{
    return ddv{p ? a.hi : b.hi, p ? a.lo : b.lo};
}
DDI zddv select_any(bool p, const zddv &a, const zddv &b)
{
    return zddv{select_any(p, a.re, b.re), select_any(p, a.im, b.im)};
}

DDI bool tri_diag_row(int64_t j, int64_t r, int64_t n, int64_t nb, bool upper, int64_t &i)
{
    const int64_t jc = (j / nb) * nb;
    if (upper) {
        i = jc + r;
        return i <= j;
    }
    i = j + r;
    return i < jc + nb && i < n;
}

// The chain over l is sequential, so operands are loaded DIAG_PF steps ahead of their use:
// each lane then waits once per DIAG_PF steps on memory instead of once per step.  A batch
// is branch-free (a zero t_l keeps c by a select, which also keeps a NaN or Inf of a
// skipped column out of c), so the compiler issues all of its loads first.
constexpr int DIAG_PF = 8;

template <bool TA, typename V>
DDI void tr_step(V &c, V &acc, const V &alpha, const V &a, const V &b)
{
    if constexpr (!TA) {
        const V tl = mul_exact_any(alpha, b);
        const V upd = add_exact_any(c, mul_exact_any(tl, a));
        c = select_any(is_zero_any(tl), c, upd);
    } else {
        acc = add_exact_any(acc, mul_exact_any(a, b));
    }
}

template <bool TA, bool TB>
DDI void mgemmtr_diag_lane(int64_t t, int64_t n, int64_t k, int64_t nb, bool upper, ddv alpha,
                           ddv beta, int beta_mode, const ddv *__restrict__ A, int64_t lda,
                           const ddv *__restrict__ B, int64_t ldb, ddv *__restrict__ C,
                           int64_t ldc)
{
    const int64_t j = t / nb, r = t % nb;
    int64_t i;
    if (j >= n || !tri_diag_row(j, r, n, nb, upper, i)) return;
    ddv c = dd_beta_prepass(C[j * ldc + i], beta, beta_mode);
    ddv acc{0.0, 0.0};
    auto a_at = [&](int64_t l) { return TA ? A[i * lda + l] : A[l * lda + i]; };
    auto b_at = [&](int64_t l) { return TB ? B[l * ldb + j] : B[j * ldb + l]; };
    int64_t l0 = 0;
    for (; l0 + DIAG_PF <= k; l0 += DIAG_PF) {
        ddv av[DIAG_PF], bv[DIAG_PF];
#pragma unroll
        for (int q = 0; q < DIAG_PF; ++q) {
            av[q] = a_at(l0 + q);
            bv[q] = b_at(l0 + q);
        }
        __builtin_amdgcn_sched_barrier(0);   // all loads of the batch before its steps
#pragma unroll
        for (int q = 0; q < DIAG_PF; ++q) tr_step<TA>(c, acc, alpha, av[q], bv[q]);
    }
    for (; l0 < k; ++l0) tr_step<TA>(c, acc, alpha, a_at(l0), b_at(l0));
    if (TA) c = dd_add_exact(c, dd_mul_exact(alpha, acc));
    C[j * ldc + i] = c;
}

template <int OPA, int OPB>
DDI void wgemmtr_diag_lane(int64_t t, int64_t n, int64_t k, int64_t nb, bool upper,
                           zddv alpha, zddv beta, int beta_mode, const zddv *__restrict__ A,
                           int64_t lda, const zddv *__restrict__ B, int64_t ldb,
                           zddv *__restrict__ C, int64_t ldc)
{
    const int64_t j = t / nb, r = t % nb;
    int64_t i;
    if (j >= n || !tri_diag_row(j, r, n, nb, upper, i)) return;
    zddv c = zdd_beta_prepass(C[j * ldc + i], beta, beta_mode);
    zddv acc{{0.0, 0.0}, {0.0, 0.0}};
    auto a_at = [&](int64_t l) {
        const zddv a = OPA == OP_N ? A[l * lda + i] : A[i * lda + l];
        return OPA == OP_C ? zdd_conj(a) : a;
    };
    auto b_at = [&](int64_t l) {
        const zddv b = OPB == OP_N ? B[j * ldb + l] : B[l * ldb + j];
        return OPB == OP_C ? zdd_conj(b) : b;
    };
    int64_t l0 = 0;
    for (; l0 + DIAG_PF <= k; l0 += DIAG_PF) {
        zddv av[DIAG_PF], bv[DIAG_PF];
#pragma unroll
        for (int q = 0; q < DIAG_PF; ++q) {
            av[q] = a_at(l0 + q);
            bv[q] = b_at(l0 + q);
        }
        __builtin_amdgcn_sched_barrier(0);   // all loads of the batch before its steps
#pragma unroll
        for (int q = 0; q < DIAG_PF; ++q) tr_step<OPA != OP_N>(c, acc, alpha, av[q], bv[q]);
    }
    for (; l0 < k; ++l0) tr_step<OPA != OP_N>(c, acc, alpha, a_at(l0), b_at(l0));
    if (OPA != OP_N) c = zdd_add_exact(c, zdd_mul_exact(alpha, acc));
    C[j * ldc + i] = c;
}

// ----------------------------------------------------------------------------------
// Real kernel.
// ----------------------------------------------------------------------------------
// gemmtr (TRI): workgroups [0, rect_wgs) take the live rectangle tiles in the compact order of
// tri_live_tile; tiles [0, split_from) run whole, and each later live tile runs as two
// BM x BN/2 halves (two workgroups), so the last, partly filled round of resident workgroups
// lasts half a tile time (SIMD order only).  Workgroups past rect_wgs take the diagonal
// triangles.  nb32: tri_nb clamped to n + 1.
template <int TM, int TN, int BK, bool TA, bool TB, int TRI = TRI_NONE, bool SCALAR = false>
__global__ __launch_bounds__(256)
void k_mgemm_plain(int64_t m, int64_t n, int64_t k, ddv alpha, ddv beta, int beta_mode,
                   const ddv *__restrict__ A, int64_t lda, const ddv *__restrict__ B,
                   int64_t ldb, ddv *__restrict__ C, int64_t ldc, int tiles_m, int rect_wgs,
                   int64_t tri_nb, int split_from, int nb32)
{
    constexpr int BM = 16 * TM, BN = 16 * TN;
    __shared__ __attribute__((aligned(16))) ddv As[BK][BM];
    __shared__ __attribute__((aligned(16))) ddv Bs[BK][BN];
    if constexpr (TRI == TRI_NONE) {
        const int64_t i0 = (int64_t)(blockIdx.x % tiles_m) * BM;
        const int64_t j0 = (int64_t)(blockIdx.x / tiles_m) * BN;
        mgemm_plain_tile<TM, TN, BK, TA, TB, TRI, SCALAR>(m, n, k, alpha, beta, beta_mode, A, lda,
                                                          B, ldb, C, ldc, i0, j0, As, Bs);
    } else {
        const int bid = (int)blockIdx.x;
        if (bid >= rect_wgs) {
            mgemmtr_diag_lane<TA, TB>((int64_t)(bid - rect_wgs) * 256 + threadIdx.x, n, k, tri_nb,
                                      TRI == TRI_UPPER, alpha, beta, beta_mode, A, lda, B, ldb,
                                      C, ldc);
            return;
        }
        int64_t i0, j0;
        if (SCALAR || TN % 2 != 0 || bid < split_from) {
            tri_live_tile<TRI, BM, BN>(bid, (int)n, nb32, tiles_m, i0, j0);
            mgemm_plain_tile<TM, TN, BK, TA, TB, TRI, SCALAR>(m, n, k, alpha, beta, beta_mode, A,
                                                              lda, B, ldb, C, ldc, i0, j0, As, Bs,
                                                              tri_nb);
        } else if constexpr (!SCALAR && TN % 2 == 0) {
            const int h = bid - split_from;
            tri_live_tile<TRI, BM, BN>(split_from + h / 2, (int)n, nb32, tiles_m, i0, j0);
            j0 += (h & 1) * (BN / 2);
            if (j0 >= n || !tri_tile_live<TRI>(i0, j0, BM, BN / 2, n, tri_nb)) return;
            mgemm_plain_tile<TM, TN / 2, BK, TA, TB, TRI, false>(
                m, n, k, alpha, beta, beta_mode, A, lda, B, ldb, C, ldc, i0, j0, As,
                *reinterpret_cast<ddv (*)[BK][BN / 2]>(&Bs[0][0]), tri_nb);
        }
    }
}

// ----------------------------------------------------------------------------------
// Complex kernel (same structure; op letters N/T/C with conjugation in the staging).
// ----------------------------------------------------------------------------------
template <int TM, int TN, int BK, int OPA, int OPB, int TRI = TRI_NONE, bool SCALAR = false>
__global__ __launch_bounds__(256)
void k_wgemm_plain(int64_t m, int64_t n, int64_t k, zddv alpha, zddv beta, int beta_mode,
                   const zddv *__restrict__ A, int64_t lda, const zddv *__restrict__ B,
                   int64_t ldb, zddv *__restrict__ C, int64_t ldc, int tiles_m, int rect_tiles,
                   int64_t tri_nb, int diag_alt, int nb32)
{
    constexpr int BM = 16 * TM, BN = 16 * TN;
    static_assert(KCHUNK % BK == 0, "BK must divide the flush period");
    __shared__ __attribute__((aligned(32))) zddv As[BK][BM];
    __shared__ __attribute__((aligned(32))) zddv Bs[BK][BN];
    if constexpr (TRI != TRI_NONE) {
        if ((int)blockIdx.x >= rect_tiles) {
            // A letter outside N/T/C: the reference's diagonal step reads it as 'T' where the
            // GEMM packers of the rectangle may read it as 'C' (wgemmtr_serial.cpp:54-57 vs
            // wgemm_serial.cpp:56-73, 81-99); diag_alt marks those operands.
            constexpr int DA = OPA == OP_C ? OP_T : OPA, DB = OPB == OP_C ? OP_T : OPB;
            const int64_t t = (int64_t)(blockIdx.x - rect_tiles) * 256 + threadIdx.x;
            const bool up = TRI == TRI_UPPER;
            switch (diag_alt) {
            case 1:
                wgemmtr_diag_lane<DA, OPB>(t, n, k, tri_nb, up, alpha, beta, beta_mode, A, lda, B,
                                           ldb, C, ldc);
                break;
            case 2:
                wgemmtr_diag_lane<OPA, DB>(t, n, k, tri_nb, up, alpha, beta, beta_mode, A, lda, B,
                                           ldb, C, ldc);
                break;
            case 3:
                wgemmtr_diag_lane<DA, DB>(t, n, k, tri_nb, up, alpha, beta, beta_mode, A, lda, B,
                                          ldb, C, ldc);
                break;
            default:
                wgemmtr_diag_lane<OPA, OPB>(t, n, k, tri_nb, up, alpha, beta, beta_mode, A, lda, B,
                                            ldb, C, ldc);
            }
            return;
        }
    }

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    // gemmtr: one workgroup per live rectangle tile, in the compact order of tri_live_tile
    int64_t i0 = (int64_t)(blockIdx.x % tiles_m) * BM;
    int64_t j0 = (int64_t)(blockIdx.x / tiles_m) * BN;
    if constexpr (TRI != TRI_NONE) tri_live_tile<TRI, BM, BN>((int)blockIdx.x, (int)n, nb32, tiles_m, i0, j0);
    if constexpr (SCALAR) {
        scalar_order_tile<TM, TN, BK, OPA != OP_N, OPB != OP_N, TRI, OPA == OP_C, OPB == OP_C>(
            m, n, k, alpha, beta, beta_mode, A, lda, B, ldb, C, ldc, i0, j0, As, Bs, tri_nb);
        return;
    }

    SlabA<zddv, BM, BK, OPA != OP_N, OPA == OP_C> sa;
    SlabB<zddv, BN, BK, OPB != OP_N, OPB == OP_C> sb;

    zddv acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = zddv{{0.0, 0.0}, {0.0, 0.0}};
    uint32_t keep = 0;   // written elements of the lane's register tile (see mgemm_plain_tile)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int64_t gi = i0 + tx + 16 * i, gj = j0 + ty + 16 * j;
            if (gi < m && gj < n && tri_rect<TRI>(gi, gj, tri_nb)) keep |= 1u << (i * TN + j);
        }

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
        for (int kk = 0; kk < kmax; ++kk) {
            zddv a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = zdd_add_fast(acc[i][j], zdd_mul_fast(a[i], b[j]));
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
                        zddv *cp = &C[gj * ldc + gi];
                        zddv c = *cp;
                        if (first) c = zdd_beta_prepass(c, beta, beta_mode);
                        *cp = zdd_add_exact(c, zdd_mul_fast(alpha, acc[i][j]));
                    }
                    acc[i][j] = zddv{{0.0, 0.0}, {0.0, 0.0}};
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

// ----------------------------------------------------------------------------------
// beta pre-pass alone (k == 0 or alpha == 0: mgemm_parallel.cpp:68-78).
// ----------------------------------------------------------------------------------
__global__ __launch_bounds__(256)
void k_mgemm_beta_only(int64_t m, int64_t n, ddv beta, int beta_mode, ddv *__restrict__ C,
                       int64_t ldc)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    int64_t i = e % m, j = e / m;
    ddv *cp = &C[j * ldc + i];
    *cp = dd_beta_prepass(*cp, beta, beta_mode);
}

__global__ __launch_bounds__(256)
void k_wgemm_beta_only(int64_t m, int64_t n, zddv beta, int beta_mode, zddv *__restrict__ C,
                       int64_t ldc)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    int64_t i = e % m, j = e / m;
    zddv *cp = &C[j * ldc + i];
    *cp = zdd_beta_prepass(*cp, beta, beta_mode);
}

// alpha == 0 or k == 0: the triangle's beta scaling alone (mgemmtr_parallel.cpp:43-52).
template <typename V>
__global__ __launch_bounds__(256)
void k_gemmtr_beta_only(int64_t n, int upper, V beta, int beta_mode, V *__restrict__ C,
                        int64_t ldc)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    if (upper ? i > j : i < j) return;
    V *cp = &C[j * ldc + i];
    if constexpr (std::is_same<V, ddv>::value) *cp = dd_beta_prepass(*cp, beta, beta_mode);
    else *cp = zdd_beta_prepass(*cp, beta, beta_mode);
}

// ----------------------------------------------------------------------------------
// Launchers.
// ----------------------------------------------------------------------------------
namespace {
constexpr int RTM = 4, RTN = 4, RBK = 16;   // real tile: 64 x 64, 16-deep slabs
constexpr int CTM = 2, CTN = 2, CBK = 16;   // complex tile: 32 x 32

// Compute units of the current device (cached per device).
int device_cus()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    if (dev < 64) {
        const int c = cache[dev].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (dev < 64) cache[dev].store(cus, std::memory_order_relaxed);
    return cus;
}

// Live rectangle tiles of a gemmtr (see tri_live_rows).
template <int TRI>
int64_t tri_live_count(int BM, int BN, int n, int nb32, int tiles_m)
{
    int64_t live = 0;
    const int tn = (n + BN - 1) / BN;
    for (int jt = 0; jt < tn; ++jt) {
        int first = 0, cnt = 0;
        tri_live_rows<TRI>(jt, BM, BN, n, nb32, tiles_m, first, cnt);
        live += cnt;
    }
    return live;
}

// Workgroups of k_mgemm_plain<TM, TN, ...> resident on the current device at once (occupancy
// query per instantiation x compute units; 0 when unknown).
template <int TM, int TN, bool TA, bool TB, int TRI, bool SCALAR>
int64_t resident_wgs()
{
    static const int per_cu = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &b, reinterpret_cast<const void *>(k_mgemm_plain<TM, TN, RBK, TA, TB, TRI, SCALAR>),
                256, 0) != hipSuccess) {
            (void)hipGetLastError();
            b = 0;
        }
        return b;
    }();
    return (int64_t)per_cu * device_cus();
}

// Modeled length of a gemmtr rectangle launch in tile times: whole rounds of `resident`
// workgroups, plus half a tile time when the last round is at most half full (its tiles run
// as halves) or a whole one when it is fuller.
inline double tri_rounds(int64_t live, int64_t resident)
{
    if (resident <= 0) return 0.0;
    const int64_t r = live % resident;
    return (double)(live / resident) + (r == 0 ? 0.0 : 2 * r <= resident ? 0.5 : 1.0);
}

template <int TM, int TN, bool TA, bool TB, int TRI = TRI_NONE, bool SCALAR = false>
hipError_t launch_real_tile(const GemmArgsR &g, hipStream_t s, int64_t tri_nb = 0)
{
    constexpr int BM = 16 * TM, BN = 16 * TN;
    const auto kern = k_mgemm_plain<TM, TN, RBK, TA, TB, TRI, SCALAR>;
    int tiles_m = (int)((g.m + BM - 1) / BM);
    int64_t tiles = (int64_t)tiles_m * ((g.n + BN - 1) / BN);
    const int64_t diag_blocks = TRI != TRI_NONE ? (g.n * tri_nb + 255) / 256 : 0;
    if (tiles + diag_blocks > INT32_MAX) return hipErrorInvalidValue;
    int64_t rect_wgs = tiles, split_from = 0;
    int nb32 = 0;
    if constexpr (TRI != TRI_NONE) {
        // one workgroup per live tile; the tiles of the last, partly filled round of resident
        // workgroups as halves when those fit in one round
        nb32 = (int)std::min<int64_t>(tri_nb, g.n + 1);
        const int64_t live = tri_live_count<TRI>(BM, BN, (int)g.n, nb32, tiles_m);
        split_from = live;
        if (!SCALAR && TN % 2 == 0) {
            const int64_t resident = resident_wgs<TM, TN, TA, TB, TRI, SCALAR>();
            const int64_t r = resident > 0 ? live % resident : 0;
            if (r > 0 && 2 * r <= resident) split_from = live - r;
        }
        rect_wgs = split_from + 2 * (live - split_from);
        if (rect_wgs + diag_blocks > INT32_MAX) return hipErrorInvalidValue;
    }
    ScopedTimer tm(TRI != TRI_NONE ? KC_GEMMTR : KC_PLAIN_REAL, s);
    hipLaunchKernelGGL(kern, dim3((unsigned)(rect_wgs + diag_blocks)), dim3(256), 0, s, g.m, g.n,
                       g.k, g.alpha, g.beta, g.beta_mode, g.A, g.lda, g.B, g.ldb, g.C, g.ldc,
                       tiles_m, (int)rect_wgs, tri_nb, (int)split_from, nb32);
    return hipGetLastError();
}

// Variant 0: 64 x 64 tiles (4 x 4 per lane); variant 1: 32 x 64 (2 x 4 per lane), which
// keeps more workgroups in flight on small products.  Without a tuned choice, variant 1 is
// used when 64 x 64 tiles would not give every CU two workgroups.  The scalar order always
// takes 64 x 64 tiles.
template <bool TA, bool TB>
hipError_t launch_real(const GemmArgsR &g, hipStream_t s)
{
    if (g.order == ORDER_SCALAR) return launch_real_tile<RTM, RTN, TA, TB, TRI_NONE, true>(g, s);
    int v = g.variant;
    if (v < 0) v = tune_lookup(TK_PLAIN_REAL, g.m, g.n, g.k);
    if (v < 0) v = ((g.m + 63) / 64) * ((g.n + 63) / 64) < 512 ? 1 : 0;
    if (v == 1) return launch_real_tile<2, 4, TA, TB>(g, s);
    return launch_real_tile<RTM, RTN, TA, TB>(g, s);
}

template <int OA, int OB, int TRI = TRI_NONE, bool SCALAR = false>
hipError_t launch_cplx_tile(const GemmArgsZ &g, hipStream_t s, int64_t tri_nb)
{
    constexpr int BM = 16 * CTM, BN = 16 * CTN;
    int tiles_m = (int)((g.m + BM - 1) / BM);
    int64_t tiles = (int64_t)tiles_m * ((g.n + BN - 1) / BN);
    const int64_t diag_blocks = TRI != TRI_NONE ? (g.n * tri_nb + 255) / 256 : 0;
    if (tiles + diag_blocks > INT32_MAX) return hipErrorInvalidValue;
    int nb32 = 0;
    if constexpr (TRI != TRI_NONE) {   // one workgroup per live tile
        nb32 = (int)std::min<int64_t>(tri_nb, g.n + 1);
        tiles = tri_live_count<TRI>(BM, BN, (int)g.n, nb32, tiles_m);
    }
    ScopedTimer tm(TRI != TRI_NONE ? KC_GEMMTR : KC_PLAIN_CPLX, s);
    hipLaunchKernelGGL((k_wgemm_plain<CTM, CTN, CBK, OA, OB, TRI, SCALAR>),
                       dim3((unsigned)(tiles + diag_blocks)), dim3(256), 0, s, g.m, g.n, g.k,
                       g.alpha, g.beta, g.beta_mode, g.A, g.lda, g.B, g.ldb, g.C, g.ldc, tiles_m,
                       (int)tiles, tri_nb, TRI != TRI_NONE ? g.diag_alt : 0, nb32);
    return hipGetLastError();
}

template <int OA, int OB, int TRI = TRI_NONE>
hipError_t launch_cplx(const GemmArgsZ &g, hipStream_t s, int64_t tri_nb = 0)
{
    if (g.order == ORDER_SCALAR) return launch_cplx_tile<OA, OB, TRI, true>(g, s, tri_nb);
    return launch_cplx_tile<OA, OB, TRI, false>(g, s, tri_nb);
}

template <int OA>
hipError_t launch_cplx_b(const GemmArgsZ &g, hipStream_t s)
{
    switch (g.opb) {
    case OP_N: return launch_cplx<OA, OP_N>(g, s);
    case OP_T: return launch_cplx<OA, OP_T>(g, s);
    default:   return launch_cplx<OA, OP_C>(g, s);
    }
}

// Rectangle tiles of the triangular-output GEMM (SIMD order): 32 x 64 (variant 1; 164 VGPRs,
// three workgroups per CU) or 64 x 64 (variant 0; two per CU).  A 64 x 64 tile does 3.5% more
// work per CU-second, but its rounds of resident workgroups are coarser: every live tile
// runs the whole depth, so the last, partly filled round costs a whole tile time (or half,
// when its tiles run as halves; a CU left with fewer workgroups runs them barely faster).
// Without a requested variant the launcher takes the tile whose modeled length (tri_rounds)
// is shorter, a 64 x 64 round costing 1.29 times a 32 x 64 round (4.70 against 3.64 ms per
// round at k = 4096): 32 x 64 at n = 2048 and 4096, 64 x 64 from n ~ 8192 on.
constexpr double kRound64Cost = 1.29;

template <bool TA, bool TB, int TRI>
int gemmtr_tile_variant(const GemmArgsR &g, int64_t tri_nb)
{
    if (g.variant == 0 || g.variant == 1) return g.variant;
    const int n = (int)g.n, nb32 = (int)std::min<int64_t>(tri_nb, g.n + 1);
    const int64_t live32 = tri_live_count<TRI>(32, 64, n, nb32, (n + 31) / 32);
    const int64_t live64 = tri_live_count<TRI>(64, 64, n, nb32, (n + 63) / 64);
    const double r32 = tri_rounds(live32, resident_wgs<2, 4, TA, TB, TRI, false>());
    const double r64 = tri_rounds(live64, resident_wgs<RTM, RTN, TA, TB, TRI, false>());
    if (r32 <= 0.0 || r64 <= 0.0) return 1;
    return r64 * kRound64Cost < r32 ? 0 : 1;
}

// gemmtr in one launch: the masked rectangle tiles, then (workgroups past the tiles) the
// diagonal triangles, one lane per element.  The two parts write disjoint elements, and the
// diagonal workgroups, dispatched last, fill the slots that the rectangle's final, partly
// filled round of tiles leaves idle.
template <bool TA, bool TB>
hipError_t launch_gemmtr_real(const GemmArgsR &g, bool upper, int64_t nb, hipStream_t s)
{
    if (g.order == ORDER_SCALAR)
        return upper ? launch_real_tile<RTM, RTN, TA, TB, TRI_UPPER, true>(g, s, nb)
                     : launch_real_tile<RTM, RTN, TA, TB, TRI_LOWER, true>(g, s, nb);
    if (upper)
        return gemmtr_tile_variant<TA, TB, TRI_UPPER>(g, nb) == 1
                   ? launch_real_tile<2, 4, TA, TB, TRI_UPPER>(g, s, nb)
                   : launch_real_tile<RTM, RTN, TA, TB, TRI_UPPER>(g, s, nb);
    return gemmtr_tile_variant<TA, TB, TRI_LOWER>(g, nb) == 1
               ? launch_real_tile<2, 4, TA, TB, TRI_LOWER>(g, s, nb)
               : launch_real_tile<RTM, RTN, TA, TB, TRI_LOWER>(g, s, nb);
}

template <int OA, int OB>
hipError_t launch_gemmtr_cplx(const GemmArgsZ &g, bool upper, int64_t nb, hipStream_t s)
{
    return upper ? launch_cplx<OA, OB, TRI_UPPER>(g, s, nb) : launch_cplx<OA, OB, TRI_LOWER>(g, s, nb);
}

template <int OA>
hipError_t launch_gemmtr_cplx_b(const GemmArgsZ &g, bool upper, int64_t nb, hipStream_t s)
{
    switch (g.opb) {
    case OP_N: return launch_gemmtr_cplx<OA, OP_N>(g, upper, nb, s);
    case OP_T: return launch_gemmtr_cplx<OA, OP_T>(g, upper, nb, s);
    default:   return launch_gemmtr_cplx<OA, OP_C>(g, upper, nb, s);
    }
}
}  // namespace

hipError_t plain_beta_only_real(const GemmArgsR &g, hipStream_t s)
{
    int64_t cnt = g.m * g.n;
    if (cnt <= 0 || g.beta_mode == BETA_ONE) return hipSuccess;
    ScopedTimer tm(KC_BETA_ONLY, s);
    hipLaunchKernelGGL(k_mgemm_beta_only, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       g.m, g.n, g.beta, g.beta_mode, g.C, g.ldc);
    return hipGetLastError();
}

hipError_t plain_beta_only_cplx(const GemmArgsZ &g, hipStream_t s)
{
    int64_t cnt = g.m * g.n;
    if (cnt <= 0 || g.beta_mode == BETA_ONE) return hipSuccess;
    ScopedTimer tm(KC_BETA_ONLY, s);
    hipLaunchKernelGGL(k_wgemm_beta_only, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       g.m, g.n, g.beta, g.beta_mode, g.C, g.ldc);
    return hipGetLastError();
}

hipError_t plain_gemm_real(const GemmArgsR &g, hipStream_t s)
{
    const bool ta = g.opa != OP_N, tb = g.opb != OP_N;
    if (!ta && !tb) return launch_real<false, false>(g, s);
    if (ta && !tb)  return launch_real<true, false>(g, s);
    if (!ta && tb)  return launch_real<false, true>(g, s);
    return launch_real<true, true>(g, s);
}

// g.m == g.n == n.  alpha == 0 or k == 0 scale the triangle only; beta == 1 then does
// nothing (mgemmtr_parallel.cpp:43-55).
hipError_t plain_gemmtr_real(const GemmArgsR &g, bool upper, int64_t nb, hipStream_t s)
{
    if (g.n <= 0) return hipSuccess;
    if (g.k == 0 || (g.alpha.hi == 0.0 && g.alpha.lo == 0.0)) {
        if (g.beta_mode == BETA_ONE) return hipSuccess;
        ScopedTimer tm(KC_BETA_ONLY, s);
        const int64_t cnt = g.n * g.n;
        hipLaunchKernelGGL((k_gemmtr_beta_only<ddv>), dim3((unsigned)((cnt + 255) / 256)),
                           dim3(256), 0, s, g.n, upper ? 1 : 0, g.beta, g.beta_mode, g.C, g.ldc);
        return hipGetLastError();
    }
    const bool ta = g.opa != OP_N, tb = g.opb != OP_N;
    if (!ta && !tb) return launch_gemmtr_real<false, false>(g, upper, nb, s);
    if (ta && !tb)  return launch_gemmtr_real<true, false>(g, upper, nb, s);
    if (!ta && tb)  return launch_gemmtr_real<false, true>(g, upper, nb, s);
    return launch_gemmtr_real<true, true>(g, upper, nb, s);
}

hipError_t plain_gemmtr_cplx(const GemmArgsZ &g, bool upper, int64_t nb, hipStream_t s)
{
    if (g.n <= 0) return hipSuccess;
    const bool a0 = g.alpha.re.hi == 0.0 && g.alpha.re.lo == 0.0 && g.alpha.im.hi == 0.0 &&
                    g.alpha.im.lo == 0.0;
    if (g.k == 0 || a0) {
        if (g.beta_mode == BETA_ONE) return hipSuccess;
        ScopedTimer tm(KC_BETA_ONLY, s);
        const int64_t cnt = g.n * g.n;
        hipLaunchKernelGGL((k_gemmtr_beta_only<zddv>), dim3((unsigned)((cnt + 255) / 256)),
                           dim3(256), 0, s, g.n, upper ? 1 : 0, g.beta, g.beta_mode, g.C, g.ldc);
        return hipGetLastError();
    }
    switch (g.opa) {
    case OP_N: return launch_gemmtr_cplx_b<OP_N>(g, upper, nb, s);
    case OP_T: return launch_gemmtr_cplx_b<OP_T>(g, upper, nb, s);
    default:   return launch_gemmtr_cplx_b<OP_C>(g, upper, nb, s);
    }
}

hipError_t plain_gemm_cplx(const GemmArgsZ &g, hipStream_t s)
{
    switch (g.opa) {
    case OP_N: return launch_cplx_b<OP_N>(g, s);
    case OP_T: return launch_cplx_b<OP_T>(g, s);
    default:   return launch_cplx_b<OP_C>(g, s);
    }
}

}  // namespace mw
