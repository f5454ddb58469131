// ozaki.hip -- exact-product double-double GEMM on the INT8 matrix cores.
//
// Method (Ozaki scheme II, Ozaki-Uchino-Imamura 2024/25, "integer modular" emulation):
//   1. Scale every row i of op(A) by 2^sigma_i and every column j of op(B) by 2^tau_j so
//      that the largest entry of the row/column lies in [2^(b-1), 2^b); round the scaled
//      double-double to an integer A'_ik, B'_kj (|A'|,|B'| <= 2^b, error <= ~1/2 ulp of the
//      integer grid).
//   2. For each modulus m_t of a table of pairwise-coprime moduli <= 256, take symmetric
//      residues A'_t = A' mod m_t, B'_t = B' mod m_t, which fit in int8.
//   3. C'_t = A'_t B'_t on the INT8 MFMA units with INT32 accumulation -- exact, since
//      k * 128^2 < 2^31 -- reduced to u_t = C'_t mod m_t.
//   4. Reconstruct C' = A'B' exactly (|C'| < M/2, M = prod m_t) from its residues: C'/M =
//      frac(sum_t u_t y_t / m_t) with the CRT weights y_t = (M/m_t)^-1 mod m_t, evaluated in
//      128-bit fixed point (limbs of round(2^128 y_t / m_t)) with exact FP64 integer
//      arithmetic, then C' = M * that as a double-double, scaled back by 2^-(sigma+tau),
//      and combined as C := beta-prepass(C0) + alpha*C' (faithful DD ops).
// Error: normwise (relative to K*max|A|*max|B|) at most ~4*2^-b from the rounding in
// step 1; b = 103 bounds it by 4e-31.  FP64 is used only for the split, the
// reconstruction and the alpha/beta update; every product runs on the matrix cores.
//
// Complex: three real products per modulus, P0 = Ar Br^T, P1 = Ai Bi^T and
// P2 = (Ar+Ai)(Br+Bi)^T (Karatsuba, exact modulo m), from which the reconstruction forms
// Re C' = P0 - P1 and Im C' = P2 - P0 - P1 modulo m; conjugation is folded into signs.
#include "ozaki_device.hpp"
#include "timer.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace mw {

// Pairwise coprime, largest first: 2^8, 3*5*17, 11*23, 251, 13*19, primes, 7*31, primes.
static const int kModuli[OZ_MAX_MOD] = {
    256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199,
    197, 193, 191, 181, 179, 173, 167, 163, 157, 151, 149, 139, 137, 131,
    127, 113, 109, 107, 103, 101, 97,  89,  83,  79,  73,  71};

// ==================================================================================
// Host: plan construction.
// ==================================================================================
namespace {

int64_t mod_pow2(int e, int64_t m)
{
    int64_t r = 1 % m;
    for (int i = 0; i < e; ++i) r = (r * 2) % m;
    return r;
}

int64_t mod_inverse(int64_t a, int64_t m)
{
    int64_t g = m, x = 0, x1 = 1, a1 = a % m;
    while (a1) {
        int64_t q = g / a1;
        int64_t t = g - q * a1; g = a1; a1 = t;
        t = x - q * x1; x = x1; x1 = t;
    }
    return ((x % m) + m) % m;   // g == 1 for coprime inputs
}

void host_two_sum(double a, double b, double &s, double &e)
{
    volatile double t = a + b;
    volatile double bv = t - a;
    volatile double av = t - bv;
    e = (a - av) + (b - bv);
    s = t;
}

// Little-endian base-2^32 big integer -> nearest-ish double-double (error ~2^-105 rel).
void big_to_dd(const std::vector<uint32_t> &x, double &hi, double &lo)
{
    hi = 0.0;
    lo = 0.0;
    for (int i = (int)x.size() - 1; i >= 0; --i) {
        double term = std::ldexp((double)x[i], 32 * i);
        double s, e;
        host_two_sum(hi, term, s, e);
        e += lo;
        host_two_sum(s, e, hi, lo);
    }
}

std::mutex g_plan_mu;
int g_target_bits = -1;
bool g_bits_explicit = false;   // set by MWGEMM_OZAKI_BITS or ozaki_set_params
int g_forced_nmod = -1;
int g_estimate = -1;

void load_target_bits()   // with g_plan_mu held
{
    if (g_target_bits >= 0) return;
    const char *e = getenv("MWGEMM_OZAKI_BITS");
    g_target_bits = e ? atoi(e) : 103;
    g_bits_explicit = e != nullptr;
    if (g_target_bits < 8 || g_target_bits > 104) {
        g_target_bits = 103;
        g_bits_explicit = false;
    }
}

}  // namespace

int ozaki_target_bits()
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    load_target_bits();
    return g_target_bits;
}

int ozaki_estimate_bits()
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    load_target_bits();
    return g_bits_explicit ? g_target_bits : 102;
}

int ozaki_estimate()
{
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        if (g_estimate < 0) {
            const char *e = getenv("MWGEMM_OZAKI_ESTIMATE");
            g_estimate = e ? atoi(e) : 1;
            if (g_estimate < 0 || g_estimate > 2) g_estimate = 1;
        }
        if (!g_estimate) return 0;
    }
    return ozaki_forced_nmod() == 0 ? g_estimate : 0;
}

void ozaki_set_estimate(int mode)
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_estimate = mode < 0 || mode > 2 ? 1 : mode;
}

bool ozaki_estimates(int64_t m, int64_t n, int64_t k)
{
    const int mode = ozaki_estimate();
    return mode == 2 || (mode == 1 && (double)m * (double)n * (double)k >= OZ_ESTIMATE_MIN_MNK);
}

namespace {
std::atomic<int> g_count_offset{0};
}
void ozaki_set_count_offset(int offset) { g_count_offset.store(offset); }
int ozaki_count_offset() { return g_count_offset.load(); }

int ozaki_forced_nmod()
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (g_forced_nmod < 0) {
        const char *e = getenv("MWGEMM_OZAKI_NMOD");
        g_forced_nmod = e ? atoi(e) : 0;
        if (g_forced_nmod < 0 || g_forced_nmod > OZ_MAX_MOD) g_forced_nmod = 0;
    }
    return g_forced_nmod;
}

void ozaki_set_params(int bits, int nmod)
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (bits == 0) {
        g_target_bits = 103;
        g_bits_explicit = false;
    } else if (bits >= 8 && bits <= 104) {
        g_target_bits = bits;
        g_bits_explicit = true;
    }
    if (nmod >= 0 && nmod <= OZ_MAX_MOD) g_forced_nmod = nmod;
}

namespace {
void fill_plan(int n, int b, double lg, OzPlan *p);
}

bool ozaki_make_plan(int64_t k_eff, int target_bits, int force_nmod, OzPlan *p)
{
    if (k_eff <= 0 || k_eff > OZ_MAX_KEFF) return false;
    const double lk = std::log2((double)k_eff);
    int n = 0;
    double lg = 0.0;
    if (force_nmod > 0) {
        n = force_nmod;
        for (int i = 0; i < n; ++i) lg += std::log2((double)kModuli[i]);
    } else {
        // M/2 > k * 2^(2b)  <=>  log2 M > 2b + log2 k + 1; the 1e-6 margin keeps |C'|/M
        // clear of 1/2 by far more than the 2^-116 error of the fixed-point reconstruction.
        const double need = 2.0 * target_bits + lk + 1.0;
        while (n < OZ_MAX_MOD && !(lg > need + 1e-6)) lg += std::log2((double)kModuli[n++]);
        if (!(lg > need + 1e-6)) return false;
    }
    int b = (int)std::floor((lg - lk - 1.0) / 2.0);
    while (b > 0 && !(lg > 2.0 * b + lk + 1.0 + 1e-6)) --b;
    if (b > 104) b = 104;          // the split keeps every intermediate below 2^53 up to b = 104
    if (b < 1) return false;
    fill_plan(n, b, lg, p);
    return true;
}

bool ozaki_make_fixed_plan(int64_t k_eff, int bits, OzPlan *p)
{
    if (k_eff <= 0 || k_eff > OZ_MAX_KEFF || bits < 8 || bits > 104) return false;
    const double need = 2.0 * bits + std::log2((double)k_eff) + 1.0;
    int n = 0;
    double lg = 0.0;
    while (n < OZ_MAX_MOD && !(lg > need + 1e-6)) lg += std::log2((double)kModuli[n++]);
    if (!(lg > need + 1e-6)) return false;
    fill_plan(n, bits, lg, p);
    return true;
}

namespace {
void fill_plan(int n, int b, double lg, OzPlan *p)
{
    p->nmod = n;
    p->bits = b;
    p->log2M = lg;
    std::vector<uint32_t> M(1, 1u);
    for (int t = 0; t < n; ++t) {
        uint64_t carry = 0;
        for (auto &limb : M) {
            uint64_t v = (uint64_t)limb * (uint64_t)kModuli[t] + carry;
            limb = (uint32_t)v;
            carry = v >> 32;
        }
        if (carry) M.push_back((uint32_t)carry);
    }
    big_to_dd(M, p->Mhi, p->Mlo);
    for (int t = 0; t < n; ++t) {
        const int64_t m = kModuli[t];
        int64_t others = 1 % m;
        for (int s = 0; s < n; ++s)
            if (s != t) others = (others * kModuli[s]) % m;
        ModConst &c = p->mc[t];
        c.m = (double)m;
        c.inv_m = 1.0 / (double)m;
        c.y = (double)mod_inverse(others, m);
        c.p1 = (double)mod_pow2(35, m);
        c.p2 = (double)mod_pow2(70, m);
        // round(y * 2^128 / m) mod 2^128 by long division of the 5-limb number y * 2^128:
        // the CRT weight y is folded into the fixed-point limbs, so the residue GEMMs store
        // plain residues v mod m (frac(((v mod m) y mod m) / m) == frac((v mod m) y / m)).
        uint32_t num[5] = {0, 0, 0, 0, (uint32_t)c.y};
        uint32_t q[5] = {0, 0, 0, 0, 0};
        uint64_t rem = 0;
        for (int i = 4; i >= 0; --i) {
            uint64_t cur = (rem << 32) | num[i];
            q[i] = (uint32_t)(cur / (uint64_t)m);
            rem = cur % (uint64_t)m;
        }
        if (2 * rem >= (uint64_t)m) {
            for (int i = 0; i < 5; ++i) {
                if (++q[i] != 0) break;
            }
        }
        for (int l = 0; l < 4; ++l) c.w[l] = (double)q[l];
    }
}
}  // namespace

bool ozaki_supported(int64_t k) { return k >= 1 && k <= OZ_MAX_KEFF; }
bool ozaki_supported_cplx(int64_t k) { return k >= 1 && 2 * k <= OZ_MAX_KEFF; }

// ==================================================================================
// Device kernels.
// ==================================================================================
__global__ __launch_bounds__(256)
void k_oz_fill_u64(unsigned long long *p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Row maxima of |hi|+|lo| (as the bit pattern of a non-negative double, so atomicMax on the
// integer orders them; a NaN or Inf dominates).  RC: entry (r,k) at X[k*ld + r];
// KC: entry (r,k) at X[r*ld + k].
template <typename T>
__global__ __launch_bounds__(256)
void k_oz_rowmax_rc(const T *__restrict__ X, int64_t ld, int64_t R, int64_t K,
                    unsigned long long *__restrict__ rmax)
{
    const int64_t r = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t k0 = (int64_t)blockIdx.y * 256 + (threadIdx.x >> 6) * 64;
    if (r >= R) return;
    double mx = 0.0;
    const int64_t kend = (k0 + 64 < K) ? k0 + 64 : K;
    for (int64_t k = k0; k < kend; ++k) mx = nan_max(mx, mag(X[k * ld + r]));
    if (mx != 0.0 || mx != mx) atomicMax(&rmax[r], (unsigned long long)__double_as_longlong(mx));
}

template <typename T>
__global__ __launch_bounds__(256)
void k_oz_rowmax_kc(const T *__restrict__ X, int64_t ld, int64_t R, int64_t K,
                    unsigned long long *__restrict__ rmax)
{
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t k0 = (int64_t)blockIdx.y * 1024;
    double mx = 0.0;
    if (r < R) {
#pragma unroll 4
        for (int s = 0; s < 16; ++s) {
            int64_t k = k0 + lane + 64 * s;
            if (k < K) mx = nan_max(mx, mag(X[r * ld + k]));
        }
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
    // NaN-aware wave maximum on the bit patterns.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if (lane == 0 && r < R && bits != 0ull) atomicMax(&rmax[r], bits);
}

// Row maxima and residue split of a row-contiguous real operand in one kernel
// (oz_rowsplit_kc_group): groups of L = kc_lanes(Kc) lanes per row, four waves per block.
// The operand is read once from memory (a row longer than 8L is read again from L2 for the
// split), against twice for the separate maxima and split kernels.
__global__ __launch_bounds__(256)
void k_oz_rowsplit_kc(const ddv *__restrict__ X, int64_t ld, int64_t R, int64_t K, int64_t Kc,
                      int L, unsigned long long *__restrict__ rmax, int bits,
                      const ModConst *__restrict__ mcs, int nmod, int8_t *__restrict__ out,
                      int64_t mod_stride, unsigned long long *__restrict__ rnorm)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r = ((int64_t)blockIdx.x * 4 + w) * (64 / L) + lane / L;
    oz_rowsplit_kc_group(X, ld, R, K, Kc, r, L, lane & (L - 1), rmax, bits, mcs, nmod, out,
                         mod_stride, rnorm);
}

// Residue split kernel of a real operand: one unit (row r, 16-deep k group kg) per lane.
// Row-contiguous input (KC): a wave covers 64 k groups of one row.  Column-contiguous input:
// a wave covers 8 rows x 8 k groups, so the 16-byte residue chunks of one row written by a
// wave form a whole 128-byte line, and each input load reads eight 128-byte segments.
// NORM: the row norm bounds (norm_q2 sums) are added into rnorm (cleared beforehand), one
// atomic per wave and row.
template <bool KC, bool NORM>
__global__ __launch_bounds__(256)
void k_oz_split(const ddv *__restrict__ X, int64_t ld, int64_t R, int64_t K, int64_t kgroups,
                const unsigned long long *__restrict__ rmax, int bits,
                const ModConst *__restrict__ mcs, int nmod, int8_t *__restrict__ out, int64_t kp,
                int64_t mod_stride, unsigned long long *__restrict__ rnorm)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t r, kg;
    if (KC) {
        kg = (int64_t)blockIdx.y * 64 + lane;
        r = (int64_t)blockIdx.x * 4 + w;
    } else {
        r = (int64_t)blockIdx.x * 32 + w * 8 + (lane >> 3);
        kg = (int64_t)blockIdx.y * 8 + (lane & 7);
    }
    unsigned long long nrm = 0ull;
    if (r < R && kg < kgroups)
        nrm = oz_split_unit<KC, NORM>(X, ld, K, rmax, bits, mcs, nmod, out, kp, mod_stride, r, kg);
    if constexpr (NORM) {
        // Lanes of one row: the whole wave (KC), or the 8 lanes sharing lane >> 3.
        for (int off = KC ? 32 : 4; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, 64);
        if ((lane & (KC ? 63 : 7)) == 0 && r < R && nrm) atomicAdd(&rnorm[r], nrm);
    }
}

// Complex residue split, three-product form: one unit (row r, 8-deep k group kq) per lane.
// Row-contiguous input (KC): a wave covers 64 k groups of one row.  Column-contiguous input:
// a wave covers 4 rows x 16 k groups, so the 8-byte residue words of one row written by a
// wave form a 128-byte line, and each input load reads four 128-byte segments.
template <bool KC>
__global__ __launch_bounds__(256)
void k_oz_split_k3(const zddv *__restrict__ X, int64_t ld, int64_t R, int64_t K, int64_t kq_n,
                   const unsigned long long *__restrict__ rmax, int bits, int sim,
                   const ModConst *__restrict__ mcs, int nmod, int8_t *__restrict__ out,
                   int64_t kp, int64_t plane)
{
    int64_t r, kq;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (KC) {
        kq = (int64_t)blockIdx.y * 64 + lane;
        r = (int64_t)blockIdx.x * 4 + w;
    } else {
        r = (int64_t)blockIdx.x * 16 + w * 4 + (lane >> 4);
        kq = (int64_t)blockIdx.y * 16 + (lane & 15);
    }
    if (r >= R || kq >= kq_n) return;
    oz_split_unit_k3<KC>(X, ld, K, rmax, bits, sim, mcs, nmod, out, kp, plane, r, kq);
}

// Batched residue GEMM kernel: blockIdx.y = batch (modulus, or modulus x part for complex).
// FOLD: depth beyond OZ_FOLD_DEPTH (the accumulators are reduced on the way).
// dyn (data-dependent modulus count, real products): the products of moduli t >= *dyn
// return at once.
template <int TM, int TN, bool FOLD>
__global__ __launch_bounds__(256, 2)
void k_oz_gemm(const int8_t *__restrict__ A, const int8_t *__restrict__ B,
               uint8_t *__restrict__ U, int64_t kp, int64_t ldu, int64_t sA, int64_t sB,
               int64_t sU, const ModConst *__restrict__ mcs, int per, int tiles_m,
               const int *__restrict__ dyn)
{
    constexpr int BM = 64 * TM, BN = 64 * TN, BK = 64;
    __shared__ __attribute__((aligned(16))) int8_t sa[2][BM * (BK + 16)];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][BN * (BK + 16)];
    const int batch = blockIdx.y;
    if (dyn && batch / per >= *dyn) return;
    int tm, tn;
    tile_order(blockIdx.x, tiles_m, (int)gridDim.x / tiles_m, tm, tn);
    const int64_t i0 = (int64_t)tm * BM;
    const int64_t j0 = (int64_t)tn * BN;
    oz_gemm_tile<TM, TN, BK, FOLD>(A + batch * sA + i0 * kp, B + batch * sB + j0 * kp,
                                   U + batch * sU, kp, ldu, mcs + batch / per, 1, 0, 0, 0, i0, j0,
                                   sa, sb);
}

// 256 x 256 residue tiles through the DMA-fed LDS ring (eight waves, one workgroup per CU).
// blockIdx.y = batch (one product: modulus, or modulus x part for complex, per products per
// modulus).  MULTI (real, per = 1): blockIdx.y = a group of G consecutive products (moduli)
// of the tile, run as one ring sequence of `total` products in all, so the ring does not
// fill and drain per product (short depths).  M16: 16 x 16 x 64 MFMAs (the single-product
// launches: 8192^3 residue GEMMs 6-7% faster than on 32 x 32 x 32); the grouped sequences
// stay on 32 x 32 x 32, since the second A fragment buffer of the 16 x 16 x 64 form takes
// the grouped kernel past 256 VGPRs (200 spilled).
template <bool FOLD, bool MULTI, bool M16>
__global__ __launch_bounds__(RING_THREADS, 1)
void k_oz_gemm_ring(const int8_t *__restrict__ A, const int8_t *__restrict__ B,
                    uint8_t *__restrict__ U, int64_t kp, int64_t ldu, int64_t sA, int64_t sB,
                    int64_t sU, const ModConst *__restrict__ mcs, int per, int tiles_m, int G,
                    int total, const int *__restrict__ dyn)
{
    __shared__ __attribute__((aligned(1024))) int8_t smem[RING_NST * RING_SLOT];
    int tm, tn;
    tile_order(blockIdx.x, tiles_m, (int)gridDim.x / tiles_m, tm, tn);
    const int64_t i0 = (int64_t)tm * 256;
    const int64_t j0 = (int64_t)tn * 256;
    if constexpr (MULTI) {
        if (dyn && *dyn < total) total = *dyn;
        const int g0 = (int)blockIdx.y * G;
        if (g0 >= total) return;
        const int gn = total - g0 < G ? total - g0 : G;
        oz_gemm_tile_ring<4, 2, 2, 4, true, false, FOLD, M16>(A + g0 * sA + i0 * kp,
                                                         B + g0 * sB + j0 * kp, U + g0 * sU, kp,
                                                         ldu, mcs + g0, gn, sA, sB, sU, i0, j0,
                                                         smem);
    } else {
        const int batch = blockIdx.y;
        if (dyn && batch / per >= *dyn) return;
        oz_gemm_tile_ring<4, 2, 2, 4, false, false, FOLD, M16>(A + batch * sA + i0 * kp,
                                                          B + batch * sB + j0 * kp,
                                                          U + batch * sU, kp, ldu,
                                                          mcs + batch / per, 1, 0, 0, 0, i0, j0,
                                                          smem);
    }
}

// The single-product ring launch with persistent workgroups (one per CU): workgroup b runs
// the items b, b + gridDim.x, ... of the launch's tiles x products item space (RingWalk), as
// the non-persistent grid would dispatch them, with the ring kept full across items.  dyn:
// products of moduli t >= *dyn are not part of the item space.  Depth up to OZ_FOLD_DEPTH.
__global__ __launch_bounds__(RING_THREADS, 1)
void k_oz_gemm_ring_persist(const int8_t *__restrict__ A, const int8_t *__restrict__ B,
                            uint8_t *__restrict__ U, int64_t kp, int64_t ldu, int64_t sA,
                            int64_t sB, int64_t sU, const ModConst *__restrict__ mcs, int per,
                            int tiles_m, int tiles, int products, const int *__restrict__ dyn)
{
    __shared__ __attribute__((aligned(1024))) int8_t smem[RING_NST * RING_SLOT];
    if (dyn && *dyn * per < products) products = *dyn * per;
    const int items = tiles * products;
    RingWalk walk;
    walk.first = (int)blockIdx.x;
    walk.stride = (int)gridDim.x;
    walk.count = walk.first < items ? (items - walk.first + walk.stride - 1) / walk.stride : 0;
    walk.tiles_m = tiles_m;
    walk.tiles = tiles;
    walk.per = per;
    if (walk.count == 0) return;
    oz_gemm_tile_ring<4, 2, 2, 4, false, false, false, true, NoRingEpilogue, RingWalk>(
        A, B, U, kp, ldu, mcs, 1, sA, sB, sU, 0, 0, smem, 0, 0, NoRingEpilogue{}, walk);
}

// Reconstruction kernel: one 32 x 32 block of C per workgroup (oz_crt_block), blocks column
// block by column block.
// With dyn (data-dependent modulus count), mcs is the first of the per-count tables
// (sets_table) and the count *dyn, its table and its set's M (sets[*dyn]) replace nmod, mcs,
// Mhi and Mlo.  (The table stays an offset from the restrict-qualified argument, so the
// compiler still reads the constants through the scalar cache: a table pointer loaded from
// memory made every weight a vector load and cost a third of the kernel's occupancy.)
template <bool CPLX>
__global__ __launch_bounds__(256)
void k_oz_crt(const uint8_t *__restrict__ U, int64_t ldu, int64_t sU, int nmod,
              const ModConst *__restrict__ mcs, double Mhi, double Mlo,
              const unsigned long long *__restrict__ rmaxA,
              const unsigned long long *__restrict__ rmaxB, int bits, int64_t m, int64_t n,
              int64_t mblocks, typename std::conditional<CPLX, zddv, ddv>::type alpha,
              typename std::conditional<CPLX, zddv, ddv>::type beta, int beta_mode,
              typename std::conditional<CPLX, zddv, ddv>::type *__restrict__ C, int64_t ldc,
              const OzSet *__restrict__ sets, const int *__restrict__ dyn)
{
    using V = typename std::conditional<CPLX, zddv, ddv>::type;
    __shared__ V stage[32][33];
    if (dyn) {
        nmod = *dyn;
        mcs += (nmod - 1) * OZ_MAX_MOD;
        Mhi = sets[nmod].Mhi;
        Mlo = sets[nmod].Mlo;
    }
    const int64_t bi = (int64_t)blockIdx.x % mblocks, bj = (int64_t)blockIdx.x / mblocks;
    oz_crt_block<CPLX>(U, ldu, sU, nmod, mcs, Mhi, Mlo, rmaxA, rmaxB, bits, m, n, alpha, beta,
                       beta_mode, C, ldc, bi, bj, stage);
}

// The modulus count kept by the last estimating call on this device (diagnostics).
__device__ int g_oz_last_nmod;

// Data-dependent modulus count (one workgroup).  With SA, SB the largest row norm bounds of
// op(A) and op(B) (norm_q2 sums), every |C'_ij| = |sum_k A'_ik B'_kj| <= ||A'_i|| ||B'_j|| <
// sqrt(SA SB) 2^(2b-32) (Cauchy-Schwarz).  The reconstruction needs M/2 > |C'|: the count is
// the smallest n < n_hi with log2 M_n > 2b - 31 + (log2 SA + log2 SB)/2 (the 1e-6 margin
// covers the roundings of the logarithms, < 1e-13), else n_hi, the worst-case count for
// |C'| <= k 2^(2b).  SA or SB = 0 (every row holds a NaN or an infinity): the output is NaN
// regardless, one modulus.
__global__ __launch_bounds__(256)
void k_oz_choose(const unsigned long long *__restrict__ nA, int64_t RA,
                 const unsigned long long *__restrict__ nB, int64_t RB, int bits, int n_hi,
                 const OzSet *__restrict__ sets, int *__restrict__ out, int offset)
{
    __shared__ unsigned long long red[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long a = 0ull, b = 0ull;
    for (int64_t i = threadIdx.x; i < RA; i += 256) a = nA[i] > a ? nA[i] : a;
    for (int64_t j = threadIdx.x; j < RB; j += 256) b = nB[j] > b ? nB[j] : b;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long oa = __shfl_xor(a, off, 64), ob = __shfl_xor(b, off, 64);
        a = oa > a ? oa : a;
        b = ob > b ? ob : b;
    }
    if (lane == 0) {
        red[0][w] = a;
        red[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 1; q < 4; ++q) {
        a = red[0][q] > a ? red[0][q] : a;
        b = red[1][q] > b ? red[1][q] : b;
    }
    int n = n_hi;
    if (a == 0ull || b == 0ull) {
        n = 1;
    } else {
        const double need = 2.0 * bits - 31.0 + 0.5 * (log2((double)a) + log2((double)b)) + 1e-6;
        for (int c = 1; c < n_hi; ++c)
            if (sets[c].log2M > need) {
                n = c;
                break;
            }
    }
    n += offset;   // test hook (ozaki_set_count_offset), 0 otherwise
    n = n < 1 ? 1 : (n > n_hi ? n_hi : n);
    *out = n;
    g_oz_last_nmod = n;
}

// ==================================================================================
// Host orchestration.
// ==================================================================================
namespace {

// A failed HIP call also leaves its error as the thread's last error: clear it, so that the
// launch check of a later call does not report it again (sticky errors persist anyway).
#define OZ_TRY(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (void)hipGetLastError();                                                   \
            return record_error((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
        }                                                                              \
    } while (0)

struct OzDevice {
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // Completion of the last eager call that used ws: a later eager call waits for it (on
    // the device) before its kernels overwrite the workspace.
    hipEvent_t ws_done = nullptr;
    bool ws_recorded = false;
    // A call captured into a graph bakes ws's address into the graph, which may be replayed
    // at any time on any stream.  Such a buffer is never freed or reused by eager calls:
    // the next eager call moves to a buffer of its own and the captured one is kept here
    // until the process ends.
    bool ws_captured = false;
    std::vector<void *> retired;
    std::map<int, ModConst *> tables;   // keyed by modulus count
    // Every count's table and OzSet (oz_sets), for the data-dependent modulus count.
    char *sets_mem = nullptr;
    // The op(A) part of the workspace (maxima, then residues, at its start) as the last real
    // call left it: its tag (0: no reuse), shape and modulus count.
    uint64_t a_tag = 0;
    int64_t a_m = 0, a_k = 0;
    int a_nmod = 0, a_bits = 0;
    bool a_cplx = false;
    bool a_est = false;   // the part holds op(A)'s row norm bounds too
};
std::mutex g_oz_mu;
std::vector<OzDevice> g_oz;

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int oz_device(OzDevice **out)
{
    int dev = 0;
    OZ_TRY(hipGetDevice(&dev));
    if ((int)g_oz.size() <= dev) g_oz.resize(dev + 1);
    *out = &g_oz[dev];
    return 0;
}

bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// The per-device workspace for a call on stream s.  Eager calls are ordered after the
// previous eager call that used it: on the same stream by stream order, on another by
// waiting (on the device) for its completion event, which is always done when one was
// recorded (a stream handle may be reused for another stream, so equal handles prove
// nothing).  Under stream capture the workspace must already be large enough (capture can
// neither allocate nor synchronise); the captured buffer then belongs to the graph (see
// OzDevice), and replays of captured calls are ordered only by how the graphs are
// launched.  Returns OZ_NO_WORKSPACE, with nothing enqueued, when the buffer cannot be
// allocated.
int oz_workspace(OzDevice *d, size_t bytes, hipStream_t s, char **base, bool *moved = nullptr)
{
    if (moved) *moved = false;
    if (!device_alloc_allowed(bytes)) return OZ_NO_WORKSPACE;
    if (stream_capturing(s)) {
        if (d->ws_bytes < bytes)
            return record_error(-4, "ozaki: the workspace would have to grow to %zu bytes during "
                                    "stream capture; run the call once outside capture first",
                                bytes);
        d->ws_captured = true;
        *base = static_cast<char *>(d->ws);
        return 0;
    }
    if (d->ws_recorded) OZ_TRY(hipStreamWaitEvent(s, d->ws_done, 0));
    if (d->ws_captured) {
        d->retired.push_back(d->ws);
        d->ws = nullptr;
        d->ws_bytes = 0;
        d->ws_captured = false;
        if (moved) *moved = true;
    }
    if (d->ws_bytes < bytes) {
        if (d->ws) {
            if (d->ws_recorded) OZ_TRY(hipEventSynchronize(d->ws_done));
            OZ_TRY(hipStreamSynchronize(s));
            OZ_TRY(hipFree(d->ws));
            d->ws = nullptr;
            d->ws_bytes = 0;
        }
        if (device_alloc(&d->ws, bytes) != hipSuccess) {
            (void)hipGetLastError();
            d->ws = nullptr;
            return OZ_NO_WORKSPACE;
        }
        d->ws_bytes = bytes;
        if (moved) *moved = true;
    }
    *base = static_cast<char *>(d->ws);
    return 0;
}

// Records the completion of the eager call just enqueued on s (after its last kernel).
int oz_workspace_done(OzDevice *d, hipStream_t s)
{
    if (stream_capturing(s)) return 0;
    if (!d->ws_done) OZ_TRY(hipEventCreateWithFlags(&d->ws_done, hipEventDisableTiming));
    OZ_TRY(hipEventRecord(d->ws_done, s));
    d->ws_recorded = true;
    return 0;
}

int oz_table(OzDevice *d, const OzPlan &p, ModConst **out)
{
    auto it = d->tables.find(p.nmod);
    if (it != d->tables.end()) {
        *out = it->second;
        return 0;
    }
    ModConst *t = nullptr;
    if (device_alloc(reinterpret_cast<void **>(&t), sizeof(ModConst) * OZ_MAX_MOD) != hipSuccess) {
        (void)hipGetLastError();
        return OZ_NO_WORKSPACE;
    }
    if (hipMemcpy(t, p.mc, sizeof(ModConst) * p.nmod, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(t);
        return record_error(-1, "ozaki: modulus table upload failed");
    }
    d->tables[p.nmod] = t;
    *out = t;
    return 0;
}

// Device images of every modulus count's constants, at base: OZ_MAX_MOD tables of
// OZ_MAX_MOD entries (table n-1: the first n moduli with that set's CRT weights), then the
// OzSet array indexed by the count (entry n points at table n-1).
constexpr size_t OZ_TABLES_BYTES = ((sizeof(ModConst) * OZ_MAX_MOD * OZ_MAX_MOD + 255) / 256) * 256;
constexpr size_t OZ_SETS_BYTES =
    OZ_TABLES_BYTES + ((sizeof(OzSet) * (OZ_MAX_MOD + 1) + 255) / 256) * 256;

const ModConst *sets_table(const char *base, int n)
{
    return reinterpret_cast<const ModConst *>(base) + (size_t)(n - 1) * OZ_MAX_MOD;
}
const OzSet *sets_array(const char *base)
{
    return reinterpret_cast<const OzSet *>(base + OZ_TABLES_BYTES);
}

// Uploads the images for device memory at base (a synchronous copy: not under capture).
int upload_sets(char *base)
{
    static ModConst tables[OZ_MAX_MOD][OZ_MAX_MOD];
    static OzSet sets[OZ_MAX_MOD + 1];
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    sets[0] = OzSet{0.0, 0.0, 0.0};
    for (int n = 1; n <= OZ_MAX_MOD; ++n) {
        OzPlan p;
        if (!ozaki_make_plan(1, 8, n, &p)) return record_error(-5, "ozaki: modulus set %d", n);
        for (int t = 0; t < n; ++t) tables[n - 1][t] = p.mc[t];
        sets[n] = OzSet{p.Mhi, p.Mlo, p.log2M};
    }
    OZ_TRY(hipMemcpy(base, tables, sizeof tables, hipMemcpyHostToDevice));
    OZ_TRY(hipMemcpy(base + OZ_TABLES_BYTES, sets, sizeof sets, hipMemcpyHostToDevice));
    return 0;
}

// The per-device constants of every modulus count (allocated and uploaded on first use).
int oz_sets(OzDevice *d, hipStream_t s, char **out)
{
    if (!d->sets_mem) {
        if (stream_capturing(s))
            return record_error(-4, "ozaki: the first exact-product call on a device uploads its "
                                    "modulus tables and must not be captured");
        void *p = nullptr;
        if (device_alloc(&p, OZ_SETS_BYTES) != hipSuccess) {
            (void)hipGetLastError();
            return OZ_NO_WORKSPACE;
        }
        if (int rc = upload_sets(static_cast<char *>(p))) {
            (void)hipFree(p);
            return rc;
        }
        d->sets_mem = static_cast<char *>(p);
    }
    *out = d->sets_mem;
    return 0;
}

// Launch the row-maximum scan for an operand whose rows are contiguous in k (KC) or not.
template <typename T>
hipError_t launch_rowmax(const T *X, int64_t ld, int64_t R, int64_t K, bool kc,
                         unsigned long long *rmax, hipStream_t s)
{
    if (R <= 0 || K <= 0) return hipSuccess;
    if (kc) {
        dim3 grid((unsigned)((R + 3) / 4), (unsigned)((K + 1023) / 1024));
        hipLaunchKernelGGL(k_oz_rowmax_kc<T>, grid, dim3(256), 0, s, X, ld, R, K, rmax);
    } else {
        dim3 grid((unsigned)((R + 63) / 64), (unsigned)((K + 255) / 256));
        hipLaunchKernelGGL(k_oz_rowmax_rc<T>, grid, dim3(256), 0, s, X, ld, R, K, rmax);
    }
    return hipGetLastError();
}

// Residues of a real operand (kc: rows contiguous in k) into out; with rnorm (cleared), the
// row norm bounds of the data-dependent modulus count.
hipError_t launch_split(const ddv *X, int64_t ld, int64_t R, int64_t K, int64_t kc_len, bool kc,
                        const unsigned long long *rmax, int bits, const ModConst *mcs, int nmod,
                        int8_t *out, int64_t kp, int64_t mod_stride, hipStream_t s,
                        unsigned long long *rnorm = nullptr)
{
    const int64_t kgroups = kc_len / 16;
    const dim3 grid = kc ? dim3((unsigned)((R + 3) / 4), (unsigned)((kgroups + 63) / 64))
                         : dim3((unsigned)((R + 31) / 32), (unsigned)((kgroups + 7) / 8));
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, X, ld, R, K, kgroups, rmax, bits, mcs,
                           nmod, out, kp, mod_stride, rnorm);
    };
    if (kc) {
        if (rnorm) go(k_oz_split<true, true>);
        else go(k_oz_split<true, false>);
    } else {
        if (rnorm) go(k_oz_split<false, true>);
        else go(k_oz_split<false, false>);
    }
    return hipGetLastError();
}

// Three-product complex residues of one operand (kc: rows contiguous in k) into out.
hipError_t launch_split_k3(const zddv *X, int64_t ld, int64_t R, int64_t K, int64_t kc_len,
                           bool kc, const unsigned long long *rmax, int bits, int sim,
                           const ModConst *mcs, int nmod, int8_t *out, int64_t kp, int64_t plane,
                           hipStream_t s)
{
    const int64_t kq_n = kc_len / 8;
    if (kc) {
        dim3 grid((unsigned)((R + 3) / 4), (unsigned)((kq_n + 63) / 64));
        hipLaunchKernelGGL(k_oz_split_k3<true>, grid, dim3(256), 0, s, X, ld, R, K, kq_n, rmax,
                           bits, sim, mcs, nmod, out, kp, plane);
    } else {
        dim3 grid((unsigned)((R + 15) / 16), (unsigned)((kq_n + 15) / 16));
        hipLaunchKernelGGL(k_oz_split_k3<false>, grid, dim3(256), 0, s, X, ld, R, K, kq_n, rmax,
                           bits, sim, mcs, nmod, out, kp, plane);
    }
    return hipGetLastError();
}

// Row maxima and residues of a row-contiguous real operand in one kernel (with rnorm, also
// the row norm bounds).
hipError_t launch_rowsplit_kc(const ddv *X, int64_t ld, int64_t R, int64_t K, int64_t Kc,
                              unsigned long long *rmax, int bits, const ModConst *mcs, int nmod,
                              int8_t *out, int64_t mod_stride, hipStream_t s,
                              unsigned long long *rnorm = nullptr)
{
    if (R <= 0) return hipSuccess;
    const int L = kc_lanes(Kc);
    const int64_t rows_per_block = 4 * (64 / L);
    hipLaunchKernelGGL(k_oz_rowsplit_kc, dim3((unsigned)((R + rows_per_block - 1) / rows_per_block)),
                       dim3(256), 0, s, X, ld, R, K, Kc, L, rmax, bits, mcs, nmod, out, mod_stride,
                       rnorm);
    return hipGetLastError();
}

template <int TM, int TN>
hipError_t launch_gemm_tile(const int8_t *A, const int8_t *B, uint8_t *U, int64_t mp, int64_t np,
                            int64_t kp, int64_t sA, int64_t sB, int64_t sU, const ModConst *mcs,
                            int per, int batches, hipStream_t s, const int *dyn)
{
    const int tiles_m = (int)(mp / (64 * TM));
    dim3 grid((unsigned)(tiles_m * (np / (64 * TN))), (unsigned)batches);
    if (kp > OZ_FOLD_DEPTH)
        hipLaunchKernelGGL((k_oz_gemm<TM, TN, true>), grid, dim3(256), 0, s, A, B, U, kp, mp, sA,
                           sB, sU, mcs, per, tiles_m, dyn);
    else
        hipLaunchKernelGGL((k_oz_gemm<TM, TN, false>), grid, dim3(256), 0, s, A, B, U, kp, mp, sA,
                           sB, sU, mcs, per, tiles_m, dyn);
    return hipGetLastError();
}

// Persistent single-product ring launches (MWGEMM_RING_PERSIST=0: one workgroup per item).
bool ring_persist()
{
    static const bool on = [] {
        const char *e = getenv("MWGEMM_RING_PERSIST");
        return !(e && atoi(e) == 0);
    }();
    return on;
}

// Compute units of the current device (cached per device).
int device_cus()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (!cache[dev]) {
        hipDeviceProp_t prop;
        cache[dev] = hipGetDeviceProperties(&prop, dev) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    return cache[dev];
}

hipError_t launch_gemm_ring(const int8_t *A, const int8_t *B, uint8_t *U, int64_t mp, int64_t np,
                            int64_t kp, int64_t sA, int64_t sB, int64_t sU, const ModConst *mcs,
                            int per, int batches, hipStream_t s, const int *dyn)
{
    const int tiles_m = (int)(mp / 256);
    // Real products of depth <= 2048 run in groups of RING_GROUP moduli per workgroup (one
    // ring sequence each): at 8192 x 8192 the residue GEMMs take 1.61 instead of 1.85 ms at
    // k = 256, 2.12 / 2.30 at 512, 3.07 / 3.27 at 1024, 4.95 / 5.02 at 2048, and no less at
    // 4096 (where workgroups drifting apart in modulus share fewer panels in the caches).
    constexpr int RING_GROUP = 8;
    if (per == 1 && batches > 1 && kp <= 2048) {
        dim3 grid((unsigned)(tiles_m * (np / 256)), (unsigned)((batches + RING_GROUP - 1) / RING_GROUP));
        hipLaunchKernelGGL((k_oz_gemm_ring<false, true, false>), grid, dim3(RING_THREADS), 0, s, A, B, U,
                           kp, mp, sA, sB, sU, mcs, per, tiles_m, RING_GROUP, batches, dyn);
        return hipGetLastError();
    }
    const int tiles = tiles_m * (int)(np / 256);
    // Persistent workgroups unless the accumulators must fold (the folding walk does not fit
    // in 256 VGPRs).
    if (ring_persist() && kp <= OZ_FOLD_DEPTH) {
        const int cus = device_cus();
        const int64_t items = (int64_t)tiles * batches;
        const dim3 pg((unsigned)(items < cus ? items : cus));
        hipLaunchKernelGGL(k_oz_gemm_ring_persist, pg, dim3(RING_THREADS), 0, s, A, B, U, kp, mp,
                           sA, sB, sU, mcs, per, tiles_m, tiles, batches, dyn);
        return hipGetLastError();
    }
    dim3 grid((unsigned)tiles, (unsigned)batches);
    if (kp > OZ_FOLD_DEPTH)
        hipLaunchKernelGGL((k_oz_gemm_ring<true, false, true>), grid, dim3(RING_THREADS), 0, s, A, B, U,
                           kp, mp, sA, sB, sU, mcs, per, tiles_m, 1, batches, dyn);
    else
        hipLaunchKernelGGL((k_oz_gemm_ring<false, false, true>), grid, dim3(RING_THREADS), 0, s, A, B, U,
                           kp, mp, sA, sB, sU, mcs, per, tiles_m, 1, batches, dyn);
    return hipGetLastError();
}

// Variants: 0 = 256 x 256 tiles through the DMA-fed LDS ring (eight waves, one workgroup per
// CU), 1 = 128 x 128, 2 = 256 x 128 and 3 = 128 x 256 (register-staged 64-deep slabs, two
// workgroups per CU so one fills the other's barrier bubbles).  Every variant needs kp to
// be a multiple of 64.  Without a tuned choice, 256 x 256 when there are enough tiles to
// give every CU two, else 128 x 128.  dyn: the data-dependent modulus count (device word;
// the products of later moduli return at once), or null.
hipError_t launch_gemm(const int8_t *A, const int8_t *B, uint8_t *U, int64_t mp, int64_t np,
                       int64_t kp, int64_t sA, int64_t sB, int64_t sU, const ModConst *mcs,
                       int per, int batches, int variant, hipStream_t s, const int *dyn = nullptr)
{
    int v = variant;
    if (v < 0) v = tune_lookup(TK_OZ_GEMM, mp, np, kp);
    if (v < 0) v = (mp / 256) * (np / 256) * (int64_t)batches >= 512 ? 0 : 1;
    switch (v) {
    case 0: return launch_gemm_ring(A, B, U, mp, np, kp, sA, sB, sU, mcs, per, batches, s, dyn);
    case 2: return launch_gemm_tile<4, 2>(A, B, U, mp, np, kp, sA, sB, sU, mcs, per, batches, s, dyn);
    case 3: return launch_gemm_tile<2, 4>(A, B, U, mp, np, kp, sA, sB, sU, mcs, per, batches, s, dyn);
    default: return launch_gemm_tile<2, 2>(A, B, U, mp, np, kp, sA, sB, sU, mcs, per, batches, s, dyn);
    }
}

hipError_t fill_zero(unsigned long long *p, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_oz_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace

namespace {
// Workspace of one real exact-product call.  K is padded to 64 (the residue tiles consume
// 64-deep steps), rows and columns to 256.  Layout: [A maxima | A norm bounds | A residues |
// B maxima | B norm bounds | B residues | reduced products | chosen modulus count]: the
// op(A) part comes first, so that it stays in place for calls with a different n.  Each
// norm-bound array directly follows its maxima (one fill clears both).
struct OzLayoutR {
    int64_t Mp, Np, Kc;
    size_t oRA, oA, oRB, oB, oU, oDyn, bytes;
};

OzLayoutR oz_layout_real(int64_t m, int64_t n, int64_t k, int nmod)
{
    OzLayoutR L;
    L.Mp = round_up(m, 256);
    L.Np = round_up(n, 256);
    L.Kc = round_up(k, 64);
    const size_t szA = (size_t)nmod * L.Mp * L.Kc, szB = (size_t)nmod * L.Np * L.Kc;
    const size_t szU = (size_t)nmod * L.Np * L.Mp;
    L.oRA = 0;
    L.oA = 2 * L.Mp * 8;
    L.oRB = L.oA + round_up(szA, 256);
    L.oB = L.oRB + 2 * L.Np * 8;
    L.oU = L.oB + round_up(szB, 256);
    L.oDyn = L.oU + round_up(szU, 256);
    L.bytes = L.oDyn + 256 + 1024;
    return L;
}

// The kernels of one real exact-product call on the workspace at base (layout L); with
// reuse_a, op(A)'s maxima and residues (and norm bounds) are already in place.  With sets
// and tables (the device's OzSet array and per-count tables), the modulus count is
// data-dependent: plan.nmod moduli are split, the norm bounds choose how many of them the
// GEMMs and the reconstruction use.
// Only kernel launches: the maxima and norm bounds are cleared by a fill kernel, nothing is
// allocated.
int oz_enqueue_real(const GemmArgsR &g, const OzPlan &plan, const ModConst *mcs, char *base,
                    const OzLayoutR &L, bool reuse_a, hipStream_t s, const OzSet *sets,
                    const ModConst *tables)
{
    const int64_t Mp = L.Mp, Np = L.Np, Kc = L.Kc;
    const int nmod = plan.nmod;
    const bool est = sets != nullptr;
    unsigned long long *rA = reinterpret_cast<unsigned long long *>(base + L.oRA);
    unsigned long long *rB = reinterpret_cast<unsigned long long *>(base + L.oRB);
    unsigned long long *nA = est ? rA + Mp : nullptr, *nB = est ? rB + Np : nullptr;
    int *dyn = est ? reinterpret_cast<int *>(base + L.oDyn) : nullptr;
    int8_t *resA = reinterpret_cast<int8_t *>(base + L.oA);
    int8_t *resB = reinterpret_cast<int8_t *>(base + L.oB);
    uint8_t *Ub = reinterpret_cast<uint8_t *>(base + L.oU);
    const bool a_kc = g.opa != OP_N;   // op(A) row i contiguous in k when A is transposed
    const bool b_kc = g.opb == OP_N;   // op(B) column j contiguous in k when B is not
    // A row-contiguous operand of depth Kc <= 512 takes its maxima and split in one kernel
    // (the row stays in registers between them; timed with the splits).  Longer rows, and
    // column-contiguous operands, take a maxima kernel, then a split kernel: the fused one
    // would read a long row twice (measured at 8192^3: 1.83 against 1.75 ms for the pair).
    const bool a_fused = a_kc && Kc <= 512, b_fused = b_kc && Kc <= 512;
    {
        ScopedTimer tm(KC_OZ_EXPONENT, s);
        if (!reuse_a && !a_fused) {
            OZ_TRY(fill_zero(rA, est ? 2 * Mp : Mp, s));
            OZ_TRY(launch_rowmax(g.A, g.lda, g.m, g.k, a_kc, rA, s));
        }
        if (!b_fused) {
            OZ_TRY(fill_zero(rB, est ? 2 * Np : Np, s));
            OZ_TRY(launch_rowmax(g.B, g.ldb, g.n, g.k, b_kc, rB, s));
        }
    }
    {
        ScopedTimer tm(KC_OZ_SPLIT, s);
        if (!reuse_a) {
            if (a_fused)
                OZ_TRY(launch_rowsplit_kc(g.A, g.lda, g.m, g.k, Kc, rA, plan.bits, mcs, nmod, resA,
                                          Mp * Kc, s, nA));
            else
                OZ_TRY(launch_split(g.A, g.lda, g.m, g.k, Kc, a_kc, rA, plan.bits, mcs, nmod, resA,
                                    Kc, Mp * Kc, s, nA));
        }
        if (b_fused)
            OZ_TRY(launch_rowsplit_kc(g.B, g.ldb, g.n, g.k, Kc, rB, plan.bits, mcs, nmod, resB,
                                      Np * Kc, s, nB));
        else
            OZ_TRY(launch_split(g.B, g.ldb, g.n, g.k, Kc, b_kc, rB, plan.bits, mcs, nmod, resB, Kc,
                                Np * Kc, s, nB));
        if (est) {
            hipLaunchKernelGGL(k_oz_choose, dim3(1), dim3(256), 0, s, nA, g.m, nB, g.n, plan.bits,
                               nmod, sets, dyn, ozaki_count_offset());
            OZ_TRY(hipGetLastError());
        }
    }
    {
        ScopedTimer tm(KC_OZ_GEMM, s);
        OZ_TRY(launch_gemm(resA, resB, Ub, Mp, Np, Kc, Mp * Kc, Np * Kc, Np * Mp, mcs, 1, nmod,
                           g.variant, s, dyn));
    }
    {
        ScopedTimer tm(KC_OZ_CRT, s);
        const int64_t mb = (g.m + 31) / 32, nb = (g.n + 31) / 32;
        if (mb * nb > (int64_t)UINT32_MAX) return record_error(-3, "ozaki: C too large");
        hipLaunchKernelGGL((k_oz_crt<false>), dim3((unsigned)(mb * nb)), dim3(256), 0, s, Ub, Mp,
                           Np * Mp, nmod, est ? tables : mcs, plan.Mhi, plan.Mlo, rA, rB,
                           plan.bits, g.m, g.n, mb, g.alpha, g.beta, g.beta_mode, g.C, g.ldc,
                           sets, (const int *)dyn);
        OZ_TRY(hipGetLastError());
    }
    return 0;
}

// The plan of a real product: with the data-dependent count, the fixed-width worst-case
// plan whose residues the split forms; else the worst-case plan with the widest b.
bool real_plan(int64_t k, bool est, OzPlan *plan)
{
    return est ? ozaki_make_fixed_plan(k, ozaki_estimate_bits(), plan)
               : ozaki_make_plan(k, ozaki_target_bits(), ozaki_forced_nmod(), plan);
}
}  // namespace

int ozaki_gemm_real(const GemmArgsR &g, hipStream_t s, uint64_t a_tag)
{
    std::lock_guard<std::mutex> lk(g_oz_mu);
    const bool est = !g.host_call && ozaki_estimates(g.m, g.n, g.k);
    OzPlan plan;
    if (!real_plan(g.k, est, &plan))
        return record_error(-5, "ozaki: contraction depth %lld out of range", (long long)g.k);
    OzDevice *d;
    if (int rc = oz_device(&d)) return rc;
    const ModConst *mcs, *tables = nullptr;
    const OzSet *sets = nullptr;
    if (est) {
        char *sm;
        if (int rc = oz_sets(d, s, &sm)) return rc;
        mcs = sets_table(sm, plan.nmod);
        sets = sets_array(sm);
        tables = sets_table(sm, 1);
    } else {
        ModConst *t;
        if (int rc = oz_table(d, plan, &t)) return rc;
        mcs = t;
    }
    const OzLayoutR L = oz_layout_real(g.m, g.n, g.k, plan.nmod);
    char *base;
    bool moved = false;
    if (int rc = oz_workspace(d, L.bytes, s, &base, &moved)) return rc;
    const bool reuse_a = a_tag != 0 && !moved && d->a_tag == a_tag && d->a_m == g.m &&
                         d->a_k == g.k && d->a_nmod == plan.nmod && d->a_bits == plan.bits &&
                         !d->a_cplx && d->a_est == est;
    d->a_tag = 0;   // set again once this call's op(A) part is enqueued
    if (int rc = oz_enqueue_real(g, plan, mcs, base, L, reuse_a, s, sets, tables)) return rc;
    d->a_tag = a_tag;
    d->a_m = g.m;
    d->a_k = g.k;
    d->a_nmod = plan.nmod;
    d->a_bits = plan.bits;
    d->a_cplx = false;
    d->a_est = est;
    return oz_workspace_done(d, s);
}

int ozaki_last_estimate(int *nmod)
{
    OZ_TRY(hipDeviceSynchronize());
    OZ_TRY(hipMemcpyFromSymbol(nmod, HIP_SYMBOL(g_oz_last_nmod), sizeof(int), 0,
                               hipMemcpyDeviceToHost));
    return 0;
}

// ==================================================================================
// The no-allocation variant of the real device entry point.
// ==================================================================================
// Device memory of the no-allocation entry point, one copy per GPU, held until the process
// ends: the constants of every modulus count (the tables and OzSet array of upload_sets,
// written once per device at set-up, read only afterwards), then the workspace of one
// exact-product call.  The host never uses the array's name as an address: set-up looks up
// each device's copy with hipGetSymbolAddress and keeps that pointer.  Every call on a GPU
// reuses the same memory (a call clears what must start at zero), so calls that overlap on
// one GPU race.
constexpr size_t NOALLOC_TABLES = OZ_SETS_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_noalloc_mu;
std::vector<char *> g_noalloc_ptr;   // per device: its copy of scratch, once set up

// Set-up on the current device (once): the address of its copy of scratch, and the
// constants of every modulus count uploaded into its head.  Must not run under stream
// capture (the upload is a synchronous copy).
int noalloc_scratch(hipStream_t s, char **out)
{
    int dev = 0;
    OZ_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc_ptr.size() <= dev) g_noalloc_ptr.resize(dev + 1, nullptr);
    if (!g_noalloc_ptr[dev]) {
        if (stream_capturing(s))
            return record_error(-4, "no-allocation gemm: the first call on a device sets up its "
                                    "scratch memory and must not be captured");
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return record_error((int)e, "no-allocation gemm: hipGetSymbolAddress(scratch) failed: %s",
                                hipGetErrorString(e));
        }
        if (int rc = upload_sets(static_cast<char *>(ptr))) return rc;
        g_noalloc_ptr[dev] = static_cast<char *>(ptr);
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}
}  // namespace

int ozaki_gemm_real_noalloc(const GemmArgsR &g, hipStream_t s, bool may_fall_back)
{
    const bool est = ozaki_estimates(g.m, g.n, g.k);
    OzPlan plan;
    if (!real_plan(g.k, est, &plan))
        return record_error(-5, "ozaki: contraction depth %lld out of range", (long long)g.k);
    const OzLayoutR L = oz_layout_real(g.m, g.n, g.k, plan.nmod);
    if (L.bytes > NOALLOC_SCRATCH_BYTES - NOALLOC_TABLES) {
        if (may_fall_back) return OZ_NO_WORKSPACE;
        return record_error(-7, "no-allocation gemm: the exact product of %lld x %lld x %lld needs "
                                "%zu bytes of workspace, more than the %zu bytes of its scratch "
                                "memory", (long long)g.m, (long long)g.n, (long long)g.k,
                            L.bytes, NOALLOC_SCRATCH_BYTES - NOALLOC_TABLES);
    }
    char *ptr = nullptr;
    if (int rc = noalloc_scratch(s, &ptr)) return rc;
    return oz_enqueue_real(g, plan, sets_table(ptr, plan.nmod), ptr + NOALLOC_TABLES, L, false, s,
                           est ? sets_array(ptr) : nullptr, sets_table(ptr, 1));
}

int ozaki_gemm_cplx(const GemmArgsZ &g, hipStream_t s, uint64_t a_tag)
{
    std::lock_guard<std::mutex> lk(g_oz_mu);
    OzPlan plan;
    if (!ozaki_make_plan(2 * g.k, ozaki_target_bits(), ozaki_forced_nmod(), &plan))
        return record_error(-5, "ozaki: contraction depth %lld out of range", (long long)g.k);
    OzDevice *d;
    if (int rc = oz_device(&d)) return rc;
    ModConst *mcs;
    if (int rc = oz_table(d, plan, &mcs)) return rc;

    // Three products per modulus (the Karatsuba form of the complex product, exact modulo m):
    // planes 3t+q of [rows][Kc] for q = 0 (re), 1 (im), 2 (re + im) of each operand, and
    // products 3t+q of [Np][Mp].  The modulus count still covers |Re|, |Im| <= 2k 2^(2b);
    // P2 itself is never reconstructed.  Layout as for real calls: [A maxima | A residues |
    // B maxima | B residues | reduced products].
    const int64_t Mp = round_up(g.m, 256), Np = round_up(g.n, 256), Kc = round_up(g.k, 64);
    const int nmod = plan.nmod;
    const size_t szA = (size_t)3 * nmod * Mp * Kc, szB = (size_t)3 * nmod * Np * Kc;
    const size_t szU = (size_t)3 * nmod * Np * Mp;
    const size_t oRA = 0, oA = round_up(Mp * 8, 256), oRB = oA + round_up(szA, 256);
    const size_t oB = oRB + round_up(Np * 8, 256), oU = oB + round_up(szB, 256);
    char *base;
    bool moved = false;
    if (int rc = oz_workspace(d, oU + szU + 1024, s, &base, &moved)) return rc;
    unsigned long long *rA = reinterpret_cast<unsigned long long *>(base + oRA);
    unsigned long long *rB = reinterpret_cast<unsigned long long *>(base + oRB);
    int8_t *resA = reinterpret_cast<int8_t *>(base + oA);
    int8_t *resB = reinterpret_cast<int8_t *>(base + oB);
    uint8_t *Ub = reinterpret_cast<uint8_t *>(base + oU);
    const bool reuse_a = a_tag != 0 && !moved && d->a_tag == a_tag && d->a_m == g.m &&
                         d->a_k == g.k && d->a_nmod == nmod && d->a_bits == plan.bits &&
                         d->a_cplx;
    d->a_tag = 0;   // set again below once this call's op(A) part is enqueued

    const bool a_kc = g.opa != OP_N, b_kc = g.opb == OP_N;
    const int sa = g.opa == OP_C ? -1 : 1, sb = g.opb == OP_C ? -1 : 1;
    {
        ScopedTimer tm(KC_OZ_EXPONENT, s);
        if (!reuse_a) {
            OZ_TRY(fill_zero(rA, Mp, s));
            OZ_TRY(launch_rowmax(g.A, g.lda, g.m, g.k, a_kc, rA, s));
        }
        OZ_TRY(fill_zero(rB, Np, s));
        OZ_TRY(launch_rowmax(g.B, g.ldb, g.n, g.k, b_kc, rB, s));
    }
    {
        ScopedTimer tm(KC_OZ_SPLIT, s);
        if (!reuse_a)
            OZ_TRY(launch_split_k3(g.A, g.lda, g.m, g.k, Kc, a_kc, rA, plan.bits, sa, mcs, nmod,
                                   resA, Kc, Mp * Kc, s));
        OZ_TRY(launch_split_k3(g.B, g.ldb, g.n, g.k, Kc, b_kc, rB, plan.bits, sb, mcs, nmod,
                               resB, Kc, Np * Kc, s));
    }
    d->a_tag = a_tag;
    d->a_m = g.m;
    d->a_k = g.k;
    d->a_nmod = nmod;
    d->a_bits = plan.bits;
    d->a_cplx = true;
    {
        ScopedTimer tm(KC_OZ_GEMM, s);
        OZ_TRY(launch_gemm(resA, resB, Ub, Mp, Np, Kc, Mp * Kc, Np * Kc, Np * Mp, mcs, 3, 3 * nmod, g.variant, s));
    }
    {
        ScopedTimer tm(KC_OZ_CRT, s);
        const int64_t mb = (g.m + 31) / 32, nb = (g.n + 31) / 32;
        if (mb * nb > (int64_t)UINT32_MAX) return record_error(-3, "ozaki: C too large");
        hipLaunchKernelGGL((k_oz_crt<true>), dim3((unsigned)(mb * nb)), dim3(256), 0, s, Ub, Mp,
                           Np * Mp, nmod, mcs, plan.Mhi, plan.Mlo, rA, rB, plan.bits, g.m, g.n,
                           mb, g.alpha, g.beta, g.beta_mode, g.C, g.ldc, (const OzSet *)nullptr,
                           (const int *)nullptr);
        OZ_TRY(hipGetLastError());
    }
    return oz_workspace_done(d, s);
}

int ozaki_residue_gemm_test(const int8_t *dA, const int8_t *dB, uint8_t *dU, int64_t mp,
                            int64_t np, int64_t kp, int mod_index, int variant, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_oz_mu);
    if (mp % 256 || np % 256 || kp % 64 || kp <= 0 || kp > OZ_MAX_KEFF || mod_index < 0 ||
        mod_index >= OZ_MAX_MOD || variant < 0 || variant >= OZ_GEMM_VARIANTS)
        return record_error(-3, "residue gemm test: bad shape or variant");
    OzPlan plan;
    if (!ozaki_make_plan(1, 8, OZ_MAX_MOD, &plan)) return record_error(-5, "plan");
    OzDevice *d;
    if (int rc = oz_device(&d)) return rc;
    ModConst *mcs;
    if (int rc = oz_table(d, plan, &mcs)) return rc;
    OZ_TRY(launch_gemm(dA, dB, dU, mp, np, kp, 0, 0, 0, mcs + mod_index, 1, 1, variant, s));
    return 0;
}

}  // namespace mw
