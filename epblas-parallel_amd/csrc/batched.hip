// batched.hip -- many independent, unevenly sized real double-double GEMMs in one
// persistent launch (plain path) or one short fixed kernel sequence (Ozaki path), captured
// once into a HIP graph by a plan and replayed with a single hipGraphLaunch.
//
// Plain: a persistent grid of (CUs x 2) workgroups pulls 64x64 output tiles from a global
// work counter; the tile list is sorted by descending depth K (longest jobs first), so the
// queue balances uneven problems.  Each tile runs exactly the single-GEMM tile routine, so
// every problem's result is bit-identical to a separate mgemm_ call.
//
// Ozaki: the row/column maxima, residue split (maxima and split in one kernel for a
// row-contiguous operand) and reconstruction run over flattened unit spaces covering all
// problems (a wave finds its problem by binary search over prefix offsets); the residue GEMMs
// run as persistent work queues over (problem, modulus group, tile), one queue per XCD.  All
// problems share one modulus plan sized for the deepest K.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/mwgemm_hip.h"
#include "ozaki_device.hpp"
#include "plain_tile.hpp"
#include "timer.hpp"

namespace mw {

int record_error(int code, const char *fmt, ...);

// A problem takes 32 x 64 tiles instead of 64 x 64 when its rows padded to 32 are fewer than
// this fraction of its rows padded to 64.  1.0 (whenever the smaller tile wastes fewer rows)
// measured best on the 2000-problem workload (4.72 ms against 4.76 ms at 0.9 and 4.95 ms
// with 64 x 64 tiles only).
constexpr double kTile32Rate = 1.0;

// Device-side descriptor of one problem of a batch.
struct BatchProb {
    const ddv *A, *B;
    ddv *C;
    int64_t lda, ldb, ldc;
    int64_t m, n, k;
    ddv alpha, beta;
    int beta_mode, active;          // active: k > 0 and alpha != 0
    // Ozaki workspace of this problem
    int8_t *resA, *resB;
    uint8_t *U;
    unsigned long long *rA, *rB;
    int64_t Mp, Np, Kc;
};

// Flattened unit spaces of the Ozaki phases: prefix offsets per problem.
struct UnitSpace {
    const int64_t *start;           // [count + 1], start[p] = first unit of problem p
    int count;
};

DDI int find_problem(const int64_t *__restrict__ start, int count, int64_t u)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The problem of the calling wave.  Every problem's unit count is padded to a multiple of
// 64, so a wave's 64 units all belong to one problem: the search runs once per wave, on a
// wave-uniform index (scalar loads), instead of once per lane.
DDI int wave_problem(const UnitSpace &us)
{
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    return find_problem(us.start, us.count, (int64_t)blockIdx.x * 256 + w * 64);
}

__global__ __launch_bounds__(256)
void k_batch_fill(unsigned long long *p, int64_t n, int *counters, int ncounters)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
    if (i < ncounters) counters[i] = 0;
}

// ---------------------------------------------------------------------------------
// Plain persistent kernel.
// ---------------------------------------------------------------------------------
// Two workgroups per CU (at most 256 VGPRs): the tile routine alone sits just below that.
// SCALAR: the reference's scalar-kernel operation order (plain_tile.hpp), as for mgemm_.
template <bool TA, bool TB, bool SCALAR>
__global__ __launch_bounds__(256, 2)
void k_batch_plain(const BatchProb *__restrict__ probs, const int4 *__restrict__ tiles,
                   int ntiles, int *__restrict__ counter)
{
    __shared__ __attribute__((aligned(16))) ddv As[16][64];
    __shared__ __attribute__((aligned(16))) ddv Bs[16][64];
    __shared__ int s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
        __syncthreads();
        // the same for the whole workgroup: the tile record and its problem load into
        // scalar registers
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t >= ntiles) return;          // every workgroup ends here once the queue drains
        const int4 tr = tiles[t];      // (problem, tile row, tile column, shape)
        const BatchProb &p = probs[tr.x];
        if constexpr (!SCALAR) {
            if (tr.w == 1) {               // 32 x 64 (active problems only)
                mgemm_plain_tile<2, 4, 16, TA, TB, TRI_NONE, false>(
                    p.m, p.n, p.k, p.alpha, p.beta, p.beta_mode, p.A, p.lda, p.B, p.ldb, p.C,
                    p.ldc, (int64_t)tr.y * 32, (int64_t)tr.z * 64,
                    *reinterpret_cast<ddv (*)[16][32]>(&As[0][0]), Bs);
                continue;
            }
        }
        const int64_t i0 = (int64_t)tr.y * 64, j0 = (int64_t)tr.z * 64;
        if (p.active) {
            mgemm_plain_tile<4, 4, 16, TA, TB, TRI_NONE, SCALAR>(p.m, p.n, p.k, p.alpha, p.beta, p.beta_mode, p.A,
                                               p.lda, p.B, p.ldb, p.C, p.ldc, i0, j0, As, Bs);
        } else {
            const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    const int64_t gi = i0 + tx + 16 * i, gj = j0 + ty + 16 * j;
                    if (gi < p.m && gj < p.n) {
                        ddv *cp = &p.C[gj * p.ldc + gi];
                        *cp = dd_beta_prepass(*cp, p.beta, p.beta_mode);
                    }
                }
        }
    }
}

// ---------------------------------------------------------------------------------
// Ozaki phases over all problems.
// ---------------------------------------------------------------------------------
// Column-contiguous operand (stored K x rows: A for op 'N', B for op 'T'; a row-contiguous
// one goes through k_batch_rowsplit_kc below).
// Row maxima: unit = (row r, 64-deep k chunk), rows fastest so lanes read adjacent rows.
template <bool A_SIDE>
__global__ __launch_bounds__(256)
void k_batch_rowmax(const BatchProb *__restrict__ probs, UnitSpace us)
{
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= us.start[us.count]) return;
    const int pi = wave_problem(us);
    const BatchProb &p = probs[pi];
    const int64_t R = A_SIDE ? p.m : p.n;
    const int64_t local = u - us.start[pi];
    const ddv *X = A_SIDE ? p.A : p.B;
    const int64_t ld = A_SIDE ? p.lda : p.ldb;
    unsigned long long *rmax = A_SIDE ? p.rA : p.rB;
    if (local >= R * ((p.k + 63) / 64)) return;   // padding units
    const int64_t r = local % R, kc = local / R;
    double mx = 0.0;
    const int64_t kend = (kc * 64 + 64 < p.k) ? kc * 64 + 64 : p.k;
    for (int64_t kk = kc * 64; kk < kend; ++kk) mx = nan_max(mx, mag(X[kk * ld + r]));
    if (mx != 0.0 || mx != mx) atomicMax(&rmax[r], (unsigned long long)__double_as_longlong(mx));
}

// Residue split: unit = (row r, 16-wide k group), in blocks of 8 rows x 8 k groups per wave
// (unit counts padded to whole blocks), so the residue chunks a wave writes for one row form
// a 128-byte line and its loads read 128-byte segments.
template <bool A_SIDE>
__global__ __launch_bounds__(256)
void k_batch_split(const BatchProb *__restrict__ probs, UnitSpace us, int bits,
                   const ModConst *__restrict__ mcs, int nmod)
{
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= us.start[us.count]) return;
    const int pi = wave_problem(us);
    const BatchProb &p = probs[pi];
    const int64_t R = A_SIDE ? p.m : p.n;
    const int64_t G = p.Kc / 16;
    const int64_t local = u - us.start[pi];
    const int64_t rb8 = (R + 7) / 8, blk = local / 64, w = local % 64;
    const int64_t r = (blk % rb8) * 8 + w / 8;
    const int64_t kg = (blk / rb8) * 8 + w % 8;
    if (r >= R || kg >= G) return;
    const int64_t rows = A_SIDE ? p.Mp : p.Np;
    oz_split_unit<false>(A_SIDE ? p.A : p.B, A_SIDE ? p.lda : p.ldb, p.k, A_SIDE ? p.rA : p.rB, bits,
                  mcs, nmod, A_SIDE ? p.resA : p.resB, p.Kc, rows * p.Kc, r, kg);
}

// Row-contiguous operand (KC): row maximum and residue split in one kernel, 64 / L rows per
// wave (oz_rowsplit_kc_group; L = kc_lanes(Kc)), so short rows (the batches' K ~ 16..512)
// fill the wave instead of leaving most of its lanes idle.
template <bool A_SIDE>
__global__ __launch_bounds__(256)
void k_batch_rowsplit_kc(const BatchProb *__restrict__ probs, UnitSpace us, int bits,
                         const ModConst *__restrict__ mcs, int nmod)
{
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= us.start[us.count]) return;
    const int pi = wave_problem(us);
    const BatchProb &p = probs[pi];
    const int64_t R = A_SIDE ? p.m : p.n;
    const int L = kc_lanes(p.Kc);
    const int lane = (int)(threadIdx.x & 63);
    const int64_t r = ((u - us.start[pi]) >> 6) * (64 / L) + lane / L;
    oz_rowsplit_kc_group(A_SIDE ? p.A : p.B, A_SIDE ? p.lda : p.ldb, R, p.k, p.Kc, r, L,
                         lane & (L - 1), A_SIDE ? p.rA : p.rB, bits, mcs, nmod,
                         A_SIDE ? p.resA : p.resB, (A_SIDE ? p.Mp : p.Np) * p.Kc);
}

// Work queues of the residue GEMMs, one per XCD.  Each XCD has its own L2.  The plan assigns
// whole problems to queues (balanced by cost), so the tiles of one problem run on one XCD at
// about the same time and read its residue panels from one L2 instead of fetching them once
// per XCD.  A workgroup's home queue is the XCD it runs on, read from the hardware register
// XCC_ID (any value it returns is a valid queue index, so the mapping can only cost reuse,
// never correctness).  A workgroup takes tiles from its home queue, then from the others in
// turn (so no queue is left behind), and returns once every queue is drained.  counters[q]
// is queue q's cursor (reset before every run); qstart[q] .. qstart[q + 1] are its entries
// of the tile list.
constexpr int XCD_QUEUES = 8;

// s_getreg_b32 of HW_REG_XCC_ID (hardware register 20), bits [3:0]: the XCD of this wave.
DDI int xcd_id()
{
    return (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & (XCD_QUEUES - 1));
}

DDI int next_queued_tile(int *__restrict__ counters, const int *__restrict__ qstart, int &visited)
{
    const int home = xcd_id();
    while (visited < XCD_QUEUES) {
        const int q = (home + visited) % XCD_QUEUES;
        const int i = atomicAdd(&counters[q], 1);
        if (i < qstart[q + 1] - qstart[q]) return qstart[q] + i;
        ++visited;
    }
    return -1;
}

// Residue GEMMs: persistent queue over (problem, first modulus, tm, tn) tiles of 64*T x 64*T,
// each running the products of `group` consecutive moduli as one slab pipeline.
template <int T>
__global__ __launch_bounds__(256, 2)
void k_batch_oz_gemm(const BatchProb *__restrict__ probs, const int4 *__restrict__ tiles,
                     const int *__restrict__ qstart, int *__restrict__ counters,
                     const ModConst *__restrict__ mcs, int nmod, int group)
{
    constexpr int BT = 64 * T;
    __shared__ __attribute__((aligned(16))) int8_t sa[2][BT * 80];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][BT * 80];
    __shared__ int s_tile;
    int visited = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = next_queued_tile(counters, qstart, visited);
        __syncthreads();
        // the same for the whole workgroup: the tile record and its problem load into
        // scalar registers
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t < 0) return;
        const int4 tr = tiles[t];      // (problem, first modulus, tm, tn)
        const BatchProb &p = probs[tr.x];
        const int64_t i0 = (int64_t)tr.z * BT, j0 = (int64_t)tr.w * BT;
        const int64_t sA = p.Mp * p.Kc, sB = p.Np * p.Kc, sU = p.Np * p.Mp;
        const int nm = nmod - tr.y < group ? nmod - tr.y : group;
        oz_gemm_tile<T, T>(p.resA + tr.y * sA + i0 * p.Kc, p.resB + tr.y * sB + j0 * p.Kc,
                           p.U + tr.y * sU, p.Kc, p.Mp, mcs + tr.y, nm, sA, sB, sU, i0, j0,
                           sa, sb);
    }
}

// The same queue through the DMA-fed LDS ring (128 x 128 tiles, four stages of 64-deep
// slabs in flight): with the short K of small problems a register-staged tile waits on
// memory at every slab, the ring keeps three slabs in flight across modulus boundaries.
// Three stages of 16 KB and three workgroups per CU (12 VGPRs of each lane spilled under the
// 168-register limit that three workgroups impose): the kernel waits on its DMA, and a third
// workgroup with 9 stages in flight per CU took the residue GEMMs from 1.35-1.39 to 1.18-1.21
// ms on the 2000-problem workload (four stages and two workgroups per CU before; two stages
// and four workgroups spilled 215 registers and took 8 ms).
constexpr int BRING_NST = 3, BRING_WGS = 3;
__global__ __launch_bounds__(256, BRING_WGS)
void k_batch_oz_ring(const BatchProb *__restrict__ probs, const int4 *__restrict__ tiles,
                     const int *__restrict__ qstart, int *__restrict__ counters,
                     const ModConst *__restrict__ mcs, int nmod, int group)
{
    __shared__ __attribute__((aligned(1024))) int8_t smem[BRING_NST * ring_slot<128>()];
    __shared__ int s_tile;
    int visited = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = next_queued_tile(counters, qstart, visited);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t < 0) return;
        const int4 tr = tiles[t];      // (problem, first modulus, tm, tn)
        const BatchProb &p = probs[tr.x];
        const int64_t i0 = (int64_t)tr.z * 128, j0 = (int64_t)tr.w * 128;
        const int64_t sA = p.Mp * p.Kc, sB = p.Np * p.Kc, sU = p.Np * p.Mp;
        const int nm = nmod - tr.y < group ? nmod - tr.y : group;
        // residue rows exist up to Mp / Np (multiples of 32), possibly short of the tile
        const int ra = (int)(p.Mp - i0 < 128 ? p.Mp - i0 : 128);
        const int rb = (int)(p.Np - j0 < 128 ? p.Np - j0 : 128);
        oz_gemm_tile_ring<2, 2, 2, 2, true, true, true, false, NoRingEpilogue, NoRingWalk, BRING_NST>(
            p.resA + tr.y * sA + i0 * p.Kc, p.resB + tr.y * sB + j0 * p.Kc, p.U + tr.y * sU, p.Kc,
            p.Mp, mcs + tr.y, nm, sA, sB, sU, i0, j0, smem, ra, rb);
    }
}

// Residue GEMMs and reconstruction fused: 64 x 64 output tiles through the DMA-fed ring (four
// waves, one 32 x 32 MFMA tile each), all moduli of a tile in one stage sequence.  After each
// modulus the wave reduces its accumulators modulo m_t and adds u_t times the weight limbs of
// round(2^128 y_t / m_t) into four exact FP64 sums per entry (the same sums the separate
// reconstruction kernel forms, so the results are identical); after the last modulus it
// reconstructs its 16 entries per lane and updates C.  The reduced products never leave the
// registers, so the U arrays (write and read-back) disappear.
__global__ __launch_bounds__(256, 2)
void k_batch_oz_fused(const BatchProb *__restrict__ probs, const int4 *__restrict__ tiles,
                      const int *__restrict__ qstart, int *__restrict__ counters,
                      const ModConst *__restrict__ mcs, int nmod, double Mhi, double Mlo,
                      int bits)
{
    __shared__ __attribute__((aligned(1024))) int8_t smem[RING_NST * ring_slot<64>()];
    __shared__ int s_tile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w & 1, wn = w >> 1;
    int visited = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = next_queued_tile(counters, qstart, visited);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t < 0) return;
        const int4 tr = tiles[t];      // (problem, 0, tm, tn)
        const BatchProb &p = probs[tr.x];
        const int64_t i0 = (int64_t)tr.z * 64, j0 = (int64_t)tr.w * 64;
        const int64_t sA = p.Mp * p.Kc, sB = p.Np * p.Kc;
        const int ra = (int)(p.Mp - i0 < 64 ? p.Mp - i0 : 64);
        const int rb = (int)(p.Np - j0 < 64 ? p.Np - j0 : 64);
        // this lane's entries: column j0 + cl, rows i0 + rl + 8g + q for entry 4g + q
        double sums[16][1][4];
#pragma unroll
        for (int e = 0; e < 16; ++e)
#pragma unroll
            for (int l = 0; l < 4; ++l) sums[e][0][l] = 0.0;
        auto crt_step = [&](int mt, v16i (&acc)[1][1]) {
            const ModConst mc = mcs[mt];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const double u = (double)crt_residue(acc[0][0][e], mc);
#pragma unroll
                for (int l = 0; l < 4; ++l) sums[e][0][l] = __builtin_fma(u, mc.w[l], sums[e][0][l]);
            }
        };
        oz_gemm_tile_ring<2, 2, 1, 1, true, true>(p.resA + i0 * p.Kc, p.resB + j0 * p.Kc,
                                                  (uint8_t *)nullptr, p.Kc, p.Mp, mcs, nmod, sA,
                                                  sB, 0, i0, j0, smem, ra, rb, crt_step);
        const int64_t j = j0 + wn * 32 + (lane & 31);
        const int64_t rl = i0 + wm * 32 + 4 * (lane >> 5);
        if (j < p.n) {
            bool badj;
            const int tau = row_sigma(p.rB[j], bits, badj);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t i = rl + 8 * (e >> 2) + (e & 3);
                if (i < p.m) {
                    bool badi;
                    const int sigma = row_sigma(p.rA[i], bits, badi);
                    oz_crt_store(&p.C[j * p.ldc + i],
                                 oz_crt_value(sums[e], p.alpha, Mhi, Mlo, -128 - sigma - tau,
                                              badi || badj),
                                 p.beta, p.beta_mode);
                }
            }
        }
    }
}

// Reconstruction: 256 units (one workgroup) per 32 x 32 block of C (oz_crt_block), blocks
// column block by column block; every problem's unit count is a multiple of 256, so a
// workgroup never spans two problems.  Inactive problems get the beta pre-pass only.
// At least five workgroups per CU (at most 102 VGPRs; it needs 80 then): the reconstruction
// waits on its loads, and 5-6 workgroups per CU took it from 0.85 to 0.68-0.69 ms on the
// 2000-problem workload (four, at 107 VGPRs, before).
__global__ __launch_bounds__(256, 5)
void k_batch_crt(const BatchProb *__restrict__ probs, UnitSpace us, int nmod,
                 const ModConst *__restrict__ mcs, double Mhi, double Mlo, int bits)
{
    __shared__ ddv stage[32][33];
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= us.start[us.count]) return;          // whole workgroups
    const int pi = wave_problem(us);
    const BatchProb &p = probs[pi];
    const int64_t blk = (u - us.start[pi]) >> 8, mb = (p.m + 31) / 32;
    const int64_t bi = blk % mb, bj = blk / mb;
    if (!p.active) {
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int idx = threadIdx.x + 256 * it;
            const int64_t i = bi * 32 + (idx & 31), j = bj * 32 + (idx >> 5);
            if (i < p.m && j < p.n) {
                ddv *cp = &p.C[j * p.ldc + i];
                *cp = dd_beta_prepass(*cp, p.beta, p.beta_mode);
            }
        }
        return;
    }
    oz_crt_block<false>(p.U, p.Mp, p.Np * p.Mp, nmod, mcs, Mhi, Mlo, p.rA, p.rB, bits, p.m,
                        p.n, p.alpha, p.beta, p.beta_mode, p.C, p.ldc, bi, bj, stage);
}

// ---------------------------------------------------------------------------------
// Host side: plans.
// ---------------------------------------------------------------------------------
namespace {

// A failed HIP call also leaves its error as the thread's last error: clear it, so that the
// launch check of a later call does not report it again (sticky errors persist anyway).
#define B_TRY(expr)                                                                    \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (void)hipGetLastError();                                                   \
            return record_error((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
        }                                                                              \
    } while (0)

int64_t rup(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

unsigned blocks_for(int64_t units) { return (unsigned)((units + 255) / 256); }

}  // namespace

struct BatchPlan {
    int path = 0;
    int order = ORDER_SIMD;             // plain path: operation order (as for mgemm_)
    int ta = 0, tb = 0;                 // 1 = transposed
    int count = 0;
    int device = 0;
    void *dmem = nullptr;               // one allocation: table, lists, workspace
    BatchProb *probs = nullptr;
    int4 *tiles = nullptr;
    int ntiles = 0;
    int *counters = nullptr;            // [0]: plain queue; [0, 8): residue-GEMM queues
    int *qstart = nullptr;              // residue-GEMM queue q: tiles qstart[q] .. qstart[q+1]
    int64_t *us_ra = nullptr, *us_rb = nullptr, *us_sa = nullptr, *us_sb = nullptr, *us_crt = nullptr;
    int64_t n_ra = 0, n_rb = 0, n_sa = 0, n_sb = 0, n_crt = 0;
    unsigned long long *rmax_base = nullptr;
    int64_t rmax_count = 0;
    int grid_persist = 0;   // persistent workgroups: two per CU
    int grid_ring = 0;      // k_batch_oz_ring: BRING_WGS per CU
    hipEvent_t done = nullptr;          // completion of the last run, and its stream: a run on
    hipStream_t done_stream = nullptr;  // another stream waits for it (the runs share the
    bool done_recorded = false;         // plan's work counters and workspace)
    int oz_tile = 128;                  // residue-GEMM tile edge (128 or 64)
    bool fused = false;                 // residue GEMMs and reconstruction in one kernel
    int mod_group = 0;                  // moduli per residue tile (all of them unless split)
    OzPlan oz{};
    ModConst *mcs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double flops = 0.0;
};

// Enqueue the batch's kernels on s (used directly, and under stream capture for the graph).
static int batch_enqueue(BatchPlan *P, hipStream_t s)
{
    const bool timed = true;
    (void)timed;
    if (P->path == MWGEMM_PATH_PLAIN) {
        ScopedTimer tm(KC_BATCH_PLAIN, s);
        hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(256), 0, s, (unsigned long long *)nullptr,
                           (int64_t)0, P->counters, 1);
        B_TRY(hipGetLastError());
        if (P->ntiles == 0) return 0;
        const dim3 g((unsigned)P->grid_persist), b(256);
        auto go = [&](auto ta, auto tb) {
            if (P->order == ORDER_SCALAR)
                hipLaunchKernelGGL((k_batch_plain<decltype(ta)::value, decltype(tb)::value, true>), g,
                                   b, 0, s, P->probs, P->tiles, P->ntiles, P->counters);
            else
                hipLaunchKernelGGL((k_batch_plain<decltype(ta)::value, decltype(tb)::value, false>), g,
                                   b, 0, s, P->probs, P->tiles, P->ntiles, P->counters);
        };
        using F = std::false_type;
        using T = std::true_type;
        if (!P->ta && !P->tb) go(F{}, F{});
        else if (P->ta && !P->tb) go(T{}, F{});
        else if (!P->ta && P->tb) go(F{}, T{});
        else go(T{}, T{});
        B_TRY(hipGetLastError());
        return 0;
    }
    const bool a_kc = P->ta != 0, b_kc = P->tb == 0;
    {
        ScopedTimer tm(KC_BATCH_OZ_SPLIT, s);
        const int64_t nfill = P->rmax_count > 1 ? P->rmax_count : 1;
        hipLaunchKernelGGL(k_batch_fill, dim3(blocks_for(nfill)), dim3(256), 0, s, P->rmax_base,
                           P->rmax_count, P->counters, XCD_QUEUES);
        B_TRY(hipGetLastError());
        UnitSpace ra{P->us_ra, P->count}, rb{P->us_rb, P->count};
        UnitSpace sa{P->us_sa, P->count}, sb{P->us_sb, P->count};
        // A row-contiguous operand takes its maxima and split in one kernel; a
        // column-contiguous one its maxima kernel, then (below) its split kernel.
        if (P->n_ra) {
            if (a_kc) hipLaunchKernelGGL((k_batch_rowsplit_kc<true>), dim3(blocks_for(P->n_ra)), dim3(256), 0, s, P->probs, ra, P->oz.bits, P->mcs, P->oz.nmod);
            else      hipLaunchKernelGGL((k_batch_rowmax<true>), dim3(blocks_for(P->n_ra)), dim3(256), 0, s, P->probs, ra);
            B_TRY(hipGetLastError());
        }
        if (P->n_rb) {
            if (b_kc) hipLaunchKernelGGL((k_batch_rowsplit_kc<false>), dim3(blocks_for(P->n_rb)), dim3(256), 0, s, P->probs, rb, P->oz.bits, P->mcs, P->oz.nmod);
            else      hipLaunchKernelGGL((k_batch_rowmax<false>), dim3(blocks_for(P->n_rb)), dim3(256), 0, s, P->probs, rb);
            B_TRY(hipGetLastError());
        }
        if (P->n_sa) {
            hipLaunchKernelGGL((k_batch_split<true>), dim3(blocks_for(P->n_sa)), dim3(256), 0, s, P->probs, sa, P->oz.bits, P->mcs, P->oz.nmod);
            B_TRY(hipGetLastError());
        }
        if (P->n_sb) {
            hipLaunchKernelGGL((k_batch_split<false>), dim3(blocks_for(P->n_sb)), dim3(256), 0, s, P->probs, sb, P->oz.bits, P->mcs, P->oz.nmod);
            B_TRY(hipGetLastError());
        }
    }
    if (P->ntiles) {
        ScopedTimer tm(KC_BATCH_OZ_GEMM, s);
        if (P->fused)
            hipLaunchKernelGGL(k_batch_oz_fused, dim3((unsigned)P->grid_persist), dim3(256), 0, s,
                               P->probs, P->tiles, P->qstart, P->counters, P->mcs, P->oz.nmod,
                               P->oz.Mhi, P->oz.Mlo, P->oz.bits);
        else if (P->oz_tile == 128)
            hipLaunchKernelGGL(k_batch_oz_ring, dim3((unsigned)P->grid_ring), dim3(256), 0, s,
                               P->probs, P->tiles, P->qstart, P->counters, P->mcs, P->oz.nmod,
                               P->mod_group);
        else
            hipLaunchKernelGGL((k_batch_oz_gemm<1>), dim3((unsigned)P->grid_persist), dim3(256), 0, s,
                               P->probs, P->tiles, P->qstart, P->counters, P->mcs, P->oz.nmod,
                               P->mod_group);
        B_TRY(hipGetLastError());
    }
    if (P->n_crt) {
        ScopedTimer tm(KC_BATCH_OZ_CRT, s);
        UnitSpace cu{P->us_crt, P->count};
        hipLaunchKernelGGL(k_batch_crt, dim3(blocks_for(P->n_crt)), dim3(256), 0, s, P->probs, cu,
                           P->oz.nmod, P->mcs, P->oz.Mhi, P->oz.Mlo, P->oz.bits);
        B_TRY(hipGetLastError());
    }
    return 0;
}

static void plan_free(BatchPlan *P)
{
    if (!P) return;
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->done) (void)hipEventDestroy(P->done);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->dmem) (void)hipFree(P->dmem);
    if (P->mcs) (void)hipFree(P->mcs);
    delete P;
}

static int plan_build(BatchPlan **out, int64_t count, char transa, char transb, const int64_t *m,
                      const int64_t *n, const int64_t *k, const mw_float64x2 *alpha,
                      const mw_float64x2 *const *dA, const int64_t *lda,
                      const mw_float64x2 *const *dB, const int64_t *ldb,
                      const mw_float64x2 *beta, mw_float64x2 *const *dC, const int64_t *ldc,
                      int path, int oz_variant = -1)
{
    *out = nullptr;
    auto letter = [](char c) { c = (char)toupper((unsigned char)c); return c == 'N' ? 0 : (c == 'T' || c == 'C') ? 1 : -1; };
    const int ta = letter(transa), tb = letter(transb);
    if (ta < 0 || tb < 0) return record_error(-2, "batched mgemm: invalid trans letter");
    if (count < 0 || count > (1 << 24)) return record_error(-3, "batched mgemm: bad count %lld", (long long)count);
    if (path != MWGEMM_PATH_PLAIN && path != MWGEMM_PATH_OZAKI && path != MWGEMM_PATH_AUTO)
        return record_error(-3, "batched mgemm: bad path");
    int64_t kmax = 0;
    double flops = 0.0;
    for (int64_t p = 0; p < count; ++p) {
        const int64_t ar = ta ? k[p] : m[p], br = tb ? n[p] : k[p];
        const bool act = k[p] > 0 && !(alpha[p].limbs[0] == 0.0 && alpha[p].limbs[1] == 0.0);
        if (m[p] < 0 || n[p] < 0 || ldc[p] < (m[p] > 1 ? m[p] : 1) ||
            (act && m[p] > 0 && n[p] > 0 && (lda[p] < (ar > 1 ? ar : 1) || ldb[p] < (br > 1 ? br : 1))))
            return record_error(-3, "batched mgemm: bad shape or leading dimension in problem %lld", (long long)p);
        if (act) kmax = std::max(kmax, k[p]);
        flops += 2.0 * (double)m[p] * (double)n[p] * (double)std::max<int64_t>(k[p], 0);
    }
    // AUTO: the autotuned choice for this batch class if one was recorded; else the exact
    // product once the batch holds enough work for its fixed phases (split, maxima,
    // reconstruction) to pay off -- at least 1e9 multiply-adds over problems of mean size
    // 96 or more (the 2000-problem batch with sizes in [16,512] runs 1.2x faster there).
    const int64_t vol = count > 0 ? (int64_t)std::cbrt(flops / 2.0 / (double)count) : 0;
    const bool auto_path = path == MWGEMM_PATH_AUTO;
    if (path == MWGEMM_PATH_AUTO) {
        const int tv = tune_lookup(TK_BATCH, count, vol, kmax);
        if (tv >= 0) {
            path = tv / 16 ? MWGEMM_PATH_OZAKI : MWGEMM_PATH_PLAIN;
            if (oz_variant < 0) oz_variant = tv % 16;
        } else {
            path = flops / 2.0 >= 1.0e9 && vol >= 96 ? MWGEMM_PATH_OZAKI : MWGEMM_PATH_PLAIN;
        }
    }
    if (path == MWGEMM_PATH_OZAKI && kmax > 0 && !ozaki_supported(kmax)) {
        if (!auto_path)
            return record_error(-5, "batched ozaki: contraction depth %lld out of range", (long long)kmax);
        path = MWGEMM_PATH_PLAIN;
    }

    BatchPlan *P = new BatchPlan();
    P->path = path;
    P->order = plain_order();
    P->ta = ta;
    P->tb = tb;
    P->count = (int)count;
    P->flops = flops;
    if (hipGetDevice(&P->device) != hipSuccess) { plan_free(P); return record_error(-1, "hipGetDevice"); }
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, P->device) == hipSuccess) cus = prop.multiProcessorCount;

    std::vector<BatchProb> host(count);
    std::vector<int4> tl;
    std::vector<int64_t> ra(count + 1, 0), rb(count + 1, 0), sa(count + 1, 0), sb(count + 1, 0),
        cr(count + 1, 0);
    const bool oz = path == MWGEMM_PATH_OZAKI;
    const bool a_kc = P->ta != 0, b_kc = P->tb == 0;   // as in batch_enqueue
    // Residue-GEMM variants: 0 = 128 x 128 ring tiles, then a reconstruction kernel (the
    // default without a tuned choice); 1 = 64 x 64 register-staged tiles, then the
    // reconstruction kernel; 2 = 64 x 64 ring tiles with the reconstruction fused.
    // MWGEMM_BATCH_RESIDUE_VARIANT overrides the tuned choice.
    if (oz_variant < 0) {
        const char *e = getenv("MWGEMM_BATCH_RESIDUE_VARIANT");
        const int ev = e ? atoi(e) : -1;
        if (ev >= 0 && ev < BATCH_OZ_VARIANTS) oz_variant = ev;
    }
    if (oz_variant < 0) {
        const int tv = tune_lookup(TK_BATCH, count, vol, kmax);
        oz_variant = tv < 0 ? 0 : tv % 16;
    }
    P->fused = oz_variant == 2;
    P->oz_tile = (oz_variant == 0) ? 128 : 64;
    const int64_t OT = P->oz_tile;
    if (oz && kmax > 0 && !ozaki_make_plan(kmax, ozaki_target_bits(), ozaki_forced_nmod(), &P->oz)) {
        plan_free(P);
        return record_error(-5, "batched ozaki: no plan for k=%lld", (long long)kmax);
    }
    const int nmod = oz ? P->oz.nmod : 0;
    size_t ws = 0;                      // Ozaki workspace bytes (offsets relative to ws start)
    int64_t rmax_n = 0;
    std::vector<size_t> offA(count), offB(count), offU(count);
    std::vector<int64_t> offR(count);
    for (int64_t p = 0; p < count; ++p) {
        BatchProb &q = host[p];
        memset(&q, 0, sizeof q);
        q.A = reinterpret_cast<const ddv *>(dA ? dA[p] : nullptr);
        q.B = reinterpret_cast<const ddv *>(dB ? dB[p] : nullptr);
        q.C = reinterpret_cast<ddv *>(dC[p]);
        q.lda = lda[p]; q.ldb = ldb[p]; q.ldc = ldc[p];
        q.m = m[p]; q.n = n[p]; q.k = std::max<int64_t>(k[p], 0);
        q.alpha = ddv{alpha[p].limbs[0], alpha[p].limbs[1]};
        q.beta = ddv{beta[p].limbs[0], beta[p].limbs[1]};
        const bool b0 = q.beta.hi == 0.0 && q.beta.lo == 0.0, b1 = q.beta.hi == 1.0 && q.beta.lo == 0.0;
        q.beta_mode = b0 ? BETA_ZERO : b1 ? BETA_ONE : BETA_SCALE;
        q.active = (q.k > 0 && !(q.alpha.hi == 0.0 && q.alpha.lo == 0.0)) ? 1 : 0;
        const bool empty = q.m <= 0 || q.n <= 0;
        if (!oz) {
            if (!empty && (q.active || q.beta_mode != BETA_ONE)) {
                // 32 x 64 tiles when they cut the padded rows (a 64 x 64 tile reuses each staged
                // operand twice as often, so it is kept otherwise).
                int shape = 0, bm = 64;
                if (q.active && P->order != ORDER_SCALAR) {
                    const double a64 = (double)rup(q.m, 64), a32 = (double)rup(q.m, 32);
                    if (a32 < kTile32Rate * a64) { shape = 1; bm = 32; }
                }
                for (int64_t tn = 0; tn < (q.n + 63) / 64; ++tn)
                    for (int64_t tm = 0; tm < (q.m + bm - 1) / bm; ++tm) tl.push_back(int4{(int)p, (int)tm, (int)tn, shape});
            }
            continue;
        }
        // the reconstruction kernel's units; with the fused kernel only inactive problems
        // need it (their beta pre-pass)
        cr[p + 1] = empty || (P->fused && q.active) ? 0 : ((q.m + 31) / 32) * ((q.n + 31) / 32) * 256;
        if (empty || !q.active) continue;
        // Residue rows are padded to 32 for the ring tiles (they clamp their loads and stores
        // to the rows that exist), to 64 for the 64-row register-staged tiles.
        const int64_t PAD = (OT == 128 || P->fused) ? 32 : OT;
        q.Mp = rup(q.m, PAD); q.Np = rup(q.n, PAD); q.Kc = rup(q.k, 64);
        offR[p] = rmax_n;
        rmax_n += q.Mp + q.Np;
        offA[p] = ws; ws += rup((int64_t)nmod * q.Mp * q.Kc, 256);
        offB[p] = ws; ws += rup((int64_t)nmod * q.Np * q.Kc, 256);
        offU[p] = ws;
        if (!P->fused) ws += rup((int64_t)nmod * q.Np * q.Mp, 256);
        // unit counts of the maxima and split kernels (layouts documented there), each
        // padded to whole waves so that a wave never spans two problems
        // (a row-contiguous operand: maxima and split in one kernel, 64 / kc_lanes rows per
        // wave; a column-contiguous one: maxima, then split)
        const int64_t nch = (q.k + 63) / 64, G = q.Kc / 16, rpw = 64 / kc_lanes(q.Kc);
        ra[p + 1] = a_kc ? (q.m + rpw - 1) / rpw * 64 : rup(q.m * nch, 64);
        rb[p + 1] = b_kc ? (q.n + rpw - 1) / rpw * 64 : rup(q.n * nch, 64);
        sa[p + 1] = a_kc ? 0 : rup(q.m, 8) * rup(G, 8);
        sb[p + 1] = b_kc ? 0 : rup(q.n, 8) * rup(G, 8);
        for (int64_t tn = 0; tn < (q.Np + OT - 1) / OT; ++tn)
            for (int64_t tm = 0; tm < (q.Mp + OT - 1) / OT; ++tm) tl.push_back(int4{(int)p, 0, (int)tm, (int)tn});
    }
    // Residue tiles run all moduli in one slab pipeline; only when that leaves fewer than
    // about four tiles per CU are the moduli split into groups (tile .y = first modulus).
    // The fused reconstruction needs every modulus of a tile in one workgroup.
    if (oz && !tl.empty() && !P->fused) {
        const int64_t want = 4 * (int64_t)cus;
        int groups = (int)std::min<int64_t>(nmod, (want + (int64_t)tl.size() - 1) / (int64_t)tl.size());
        groups = std::max(groups, 1);
        P->mod_group = (nmod + groups - 1) / groups;
        if (P->mod_group < nmod) {
            std::vector<int4> split;
            split.reserve(tl.size() * (size_t)groups);
            for (const int4 &t : tl)
                for (int t0 = 0; t0 < nmod; t0 += P->mod_group) split.push_back(int4{t.x, t0, t.z, t.w});
            tl.swap(split);
        }
    }
    // Longest jobs first: plain tiles cost ~K.  Residue tiles go to one queue per XCD: whole
    // problems, the costliest first, each to the queue with the least work so far (tiles
    // x Kc), its tiles together.
    std::vector<int> qstart(XCD_QUEUES + 1, 0);
    if (!oz) {
        auto cost = [&](const int4 &t) { return host[t.x].active ? host[t.x].k * (t.w == 1 ? 1 : 2) : 0; };
        std::stable_sort(tl.begin(), tl.end(), [&](const int4 &a, const int4 &b) { return cost(a) > cost(b); });
    } else if (!tl.empty()) {
        std::vector<std::vector<int4>> per(count);
        for (const int4 &t : tl) per[t.x].push_back(t);
        std::vector<int> order;
        for (int64_t p = 0; p < count; ++p)
            if (!per[p].empty()) order.push_back((int)p);
        auto pcost = [&](int p) { return (double)per[p].size() * (double)host[p].Kc; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pcost(a) > pcost(b); });
        std::vector<std::vector<int4>> queue(XCD_QUEUES);
        std::vector<double> load(XCD_QUEUES, 0.0);
        for (int p : order) {
            const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[q] += pcost(p);
            queue[q].insert(queue[q].end(), per[p].begin(), per[p].end());
        }
        tl.clear();
        for (int q = 0; q < XCD_QUEUES; ++q) {
            qstart[q] = (int)tl.size();
            tl.insert(tl.end(), queue[q].begin(), queue[q].end());
        }
    }
    qstart[XCD_QUEUES] = (int)tl.size();
    for (int64_t p = 0; p < count; ++p) {
        ra[p + 1] += ra[p]; rb[p + 1] += rb[p]; sa[p + 1] += sa[p]; sb[p + 1] += sb[p]; cr[p + 1] += cr[p];
    }
    if (tl.size() > (size_t)INT32_MAX) { plan_free(P); return record_error(-3, "batched mgemm: too many tiles"); }
    P->ntiles = (int)tl.size();
    P->n_ra = ra[count]; P->n_rb = rb[count]; P->n_sa = sa[count]; P->n_sb = sb[count]; P->n_crt = cr[count];
    P->rmax_count = rmax_n;
    P->grid_persist = std::max(1, std::min<int>(P->ntiles, cus * 2));
    P->grid_ring = std::max(1, std::min<int>(P->ntiles, cus * BRING_WGS));

    // One device allocation: [probs | tiles | counters | 5 prefix arrays | rmax | workspace]
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_probs = take(sizeof(BatchProb) * std::max<int64_t>(count, 1));
    const size_t o_tiles = take(sizeof(int4) * std::max<size_t>(tl.size(), 1));
    const size_t o_cnt = take(sizeof(int) * 64);
    const size_t o_q = take(sizeof(int) * (XCD_QUEUES + 1));
    const size_t pre = sizeof(int64_t) * (count + 1);
    const size_t o_ra = take(pre), o_rb = take(pre), o_sa = take(pre), o_sb = take(pre), o_cr = take(pre);
    const size_t o_rmax = take(sizeof(unsigned long long) * std::max<int64_t>(rmax_n, 1));
    const size_t o_ws = take(std::max<size_t>(ws, 1));
    if (device_alloc(&P->dmem, off) != hipSuccess) {
        (void)hipGetLastError();
        P->dmem = nullptr;
        size_t bytes = off;
        plan_free(P);
        // Without its workspace the exact product runs on the plain path, whose plan holds
        // only the problem and tile tables (the reference's driver likewise recomputes
        // when its pack buffers cannot be had, mgemm_parallel.cpp:180-186).
        if (oz)
            return plan_build(out, count, transa, transb, m, n, k, alpha, dA, lda, dB, ldb, beta,
                              dC, ldc, MWGEMM_PATH_PLAIN);
        return record_error(-6, "batched mgemm: cannot allocate %zu bytes of plan memory", bytes);
    }
    char *base = static_cast<char *>(P->dmem);
    P->probs = reinterpret_cast<BatchProb *>(base + o_probs);
    P->tiles = reinterpret_cast<int4 *>(base + o_tiles);
    P->counters = reinterpret_cast<int *>(base + o_cnt);
    P->qstart = reinterpret_cast<int *>(base + o_q);
    P->us_ra = reinterpret_cast<int64_t *>(base + o_ra);
    P->us_rb = reinterpret_cast<int64_t *>(base + o_rb);
    P->us_sa = reinterpret_cast<int64_t *>(base + o_sa);
    P->us_sb = reinterpret_cast<int64_t *>(base + o_sb);
    P->us_crt = reinterpret_cast<int64_t *>(base + o_cr);
    P->rmax_base = reinterpret_cast<unsigned long long *>(base + o_rmax);
    char *wsb = base + o_ws;
    for (int64_t p = 0; p < count; ++p) {
        BatchProb &q = host[p];
        if (!oz || q.Kc == 0) continue;
        q.rA = P->rmax_base + offR[p];
        q.rB = q.rA + q.Mp;
        q.resA = reinterpret_cast<int8_t *>(wsb + offA[p]);
        q.resB = reinterpret_cast<int8_t *>(wsb + offB[p]);
        q.U = reinterpret_cast<uint8_t *>(wsb + offU[p]);
    }
    int rc = 0;
    if (hipMemcpy(P->probs, host.data(), sizeof(BatchProb) * count, hipMemcpyHostToDevice) != hipSuccess ||
        (!tl.empty() && hipMemcpy(P->tiles, tl.data(), sizeof(int4) * tl.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(P->us_ra, ra.data(), pre, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->us_rb, rb.data(), pre, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->us_sa, sa.data(), pre, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->us_sb, sb.data(), pre, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->us_crt, cr.data(), pre, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->qstart, qstart.data(), sizeof(int) * (XCD_QUEUES + 1), hipMemcpyHostToDevice) != hipSuccess)
    {
        (void)hipGetLastError();
        rc = record_error(-1, "batched mgemm: plan upload failed");
    }
    if (!rc && oz && nmod > 0) {
        if (device_alloc(reinterpret_cast<void **>(&P->mcs), sizeof(ModConst) * OZ_MAX_MOD) != hipSuccess ||
            hipMemcpy(P->mcs, P->oz.mc, sizeof(ModConst) * nmod, hipMemcpyHostToDevice) != hipSuccess)
        {
            (void)hipGetLastError();
            rc = record_error(-1, "batched mgemm: modulus table upload failed");
        }
    }
    if (rc) { plan_free(P); return rc; }
    *out = P;
    return 0;
}

// Autotuner hook: build a plan (no graph) with an explicit path / residue-tile variant,
// run it reps times on s and return the best time in milliseconds (or a negative code).
double batch_time_variant(int64_t count, char ta, char tb, const int64_t *m, const int64_t *n,
                          const int64_t *k, const mw_float64x2 *alpha,
                          const mw_float64x2 *const *dA, const int64_t *lda,
                          const mw_float64x2 *const *dB, const int64_t *ldb,
                          const mw_float64x2 *beta, mw_float64x2 *const *dC, const int64_t *ldc,
                          int path, int oz_variant, int reps, hipStream_t s)
{
    BatchPlan *P = nullptr;
    if (int rc = plan_build(&P, count, ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, path, oz_variant))
        return (double)(rc < 0 ? rc : -rc);
    hipEvent_t e0, e1;
    double best = -1.0;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        if (batch_enqueue(P, s) == 0 && hipStreamSynchronize(s) == hipSuccess) {
            for (int r = 0; r < reps; ++r) {
                float ms = 0.f;
                if (hipEventRecord(e0, s) != hipSuccess || batch_enqueue(P, s) != 0 ||
                    hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    break;
                if (best < 0 || ms < best) best = ms;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipStreamSynchronize(s);
    plan_free(P);
    return best;
}

}  // namespace mw

using namespace mw;

struct mwgemm_batched_plan_s {
    BatchPlan *p;
};

extern "C" {

int mwgemm_batched_plan_create(mwgemm_batched_plan *plan, int64_t count, char transa, char transb,
                               const int64_t *m, const int64_t *n, const int64_t *k,
                               const mw_float64x2 *alpha, const mw_float64x2 *const *dA,
                               const int64_t *lda, const mw_float64x2 *const *dB,
                               const int64_t *ldb, const mw_float64x2 *beta,
                               mw_float64x2 *const *dC, const int64_t *ldc, int path)
{
    if (!plan) return record_error(-3, "batched plan: null handle");
    *plan = nullptr;
    BatchPlan *P = nullptr;
    if (int rc = plan_build(&P, count, transa, transb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, path))
        return rc;
    // Capture the kernel sequence once; runs replay it with one graph launch.
    hipStream_t cs = nullptr;
    int rc = 0;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) rc = record_error(-1, "stream");
    if (!rc && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = record_error(-1, "begin capture");
    if (!rc) {
        int erc = batch_enqueue(P, cs);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (erc) rc = erc;
        else if (e != hipSuccess) rc = record_error((int)e, "end capture: %s", hipGetErrorString(e));
        else {
            P->graph = g;
            e = hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = record_error((int)e, "graph instantiate: %s", hipGetErrorString(e));
        }
    }
    if (cs) (void)hipStreamDestroy(cs);
    if (rc) { plan_free(P); return rc; }
    auto *h = new mwgemm_batched_plan_s{P};
    *plan = h;
    return 0;
}

int mwgemm_batched_plan_run(mwgemm_batched_plan plan, void *stream)
{
    if (!plan || !plan->p || !plan->p->exec) return record_error(-3, "batched plan: invalid handle");
    BatchPlan *P = plan->p;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const bool capturing = hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
    // Always wait when an earlier run was recorded: a stream handle may be reused for another
    // stream, so an equal handle proves nothing (the wait is free on the same stream).
    if (!capturing && P->done_recorded) {
        hipError_t w = hipStreamWaitEvent(s, P->done, 0);
        if (w != hipSuccess) return record_error((int)w, "hipStreamWaitEvent: %s", hipGetErrorString(w));
    }
    hipError_t e = hipGraphLaunch(P->exec, s);
    if (e != hipSuccess) return record_error((int)e, "hipGraphLaunch: %s", hipGetErrorString(e));
    if (!capturing) {
        if (!P->done && (e = hipEventCreateWithFlags(&P->done, hipEventDisableTiming)) != hipSuccess)
            return record_error((int)e, "hipEventCreate: %s", hipGetErrorString(e));
        if ((e = hipEventRecord(P->done, s)) != hipSuccess)
            return record_error((int)e, "hipEventRecord: %s", hipGetErrorString(e));
        P->done_stream = s;
        P->done_recorded = true;
    }
    return 0;
}

int mwgemm_batched_plan_info(mwgemm_batched_plan plan, int *path, int64_t *tiles, double *flops,
                             int *nmod)
{
    if (!plan || !plan->p) return -1;
    if (path) *path = plan->p->path;
    if (tiles) *tiles = plan->p->ntiles;
    if (flops) *flops = plan->p->flops;
    if (nmod) *nmod = plan->p->path == MWGEMM_PATH_OZAKI ? plan->p->oz.nmod : 0;
    return 0;
}

int mwgemm_batched_plan_destroy(mwgemm_batched_plan plan)
{
    if (!plan) return 0;
    (void)hipDeviceSynchronize();
    plan_free(plan->p);
    delete plan;
    return 0;
}

int mwgemm_mgemm_batched_device(int64_t count, char transa, char transb, const int64_t *m,
                                const int64_t *n, const int64_t *k, const mw_float64x2 *alpha,
                                const mw_float64x2 *const *dA, const int64_t *lda,
                                const mw_float64x2 *const *dB, const int64_t *ldb,
                                const mw_float64x2 *beta, mw_float64x2 *const *dC,
                                const int64_t *ldc, int path, void *stream)
{
    BatchPlan *P = nullptr;
    if (int rc = plan_build(&P, count, transa, transb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, path))
        return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = batch_enqueue(P, s);
    // The plan's memory is released once the stream has drained it.
    hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = record_error((int)e, "batched mgemm: %s", hipGetErrorString(e));
    plan_free(P);
    return rc;
}

}  // extern "C"
