// ozaki_device.hpp -- device-side building blocks of the exact-product (Ozaki scheme II)
// double-double GEMM: integer chunking and residues, the INT8 MFMA residue-GEMM tile, and
// the fixed-point Chinese-remainder reconstruction.  Shared by the single-GEMM kernels
// (ozaki.hip) and the persistent batched kernels (batched.hip).
#pragma once
#include <type_traits>
#include "ozaki.hpp"

namespace mw {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Maximum that keeps a NaN (fmax would drop it): a NaN entry must poison its row.
DDI double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }
DDI double mag(const ddv &x) { return fabs(x.hi) + fabs(x.lo); }
DDI double mag(const zddv &x) { return nan_max(mag(x.re), mag(x.im)); }

// Scale exponent of a row from its maximum: largest entry lands in [2^(b-1), 2^b).
DDI int row_sigma(unsigned long long bits, int b, bool &bad)
{
    double mx = __longlong_as_double((long long)bits);
    bad = !(mx <= 1.79769313486231570815e308);   // NaN or Inf
    if (bad || mx == 0.0) return 0;
    int e;
    (void)frexp(mx, &e);
    return b - e;
}

// The integer round(x * 2^sigma) of a DD value as three exact FP64 chunks with
// value = c0 + c1*2^35 + c2*2^70, 0 <= c0 < 2^35, |c1| < 2^36, |c2| <= 2^34 (b <= 104).
// The chunks are cut from |xh| and then signed (the low bits of a 53-bit significand are
// always representable, whereas xh mod 2^70 of a negative xh generally is not); the low
// integer xl is then folded into c0 with a carry into c1.
DDI void int_chunks(double hi, double lo, int sigma, double &c0, double &c1, double &c2)
{
    const double xh_s = ldexp(hi, sigma), xl_s = ldexp(lo, sigma);
    const double xh = rint(xh_s);
    const double xl = rint((xh_s - xh) + xl_s);
    const double ax = fabs(xh);
    const double a2 = floor(ax * 0x1p-70);
    const double rem = __builtin_fma(-a2, 0x1p70, ax);
    const double a1 = floor(rem * 0x1p-35);
    const double a0 = __builtin_fma(-a1, 0x1p35, rem);
    const double sg = xh < 0.0 ? -1.0 : 1.0;
    const double t0 = __builtin_fma(sg, a0, xl);          // |t0| < 2^52, exact
    const double carry = floor(t0 * 0x1p-35);
    c0 = __builtin_fma(-carry, 0x1p35, t0);
    c1 = __builtin_fma(sg, a1, carry);
    c2 = sg * a2;
}

// Data-dependent modulus count: a bound on the Euclidean norm of a scaled row.  An entry x
// of a row scaled by 2^sigma becomes the integer X' of int_chunks, within 5/8 of
// 2^sigma (hi + lo), so |X'| < 2^sigma (|hi| + |lo|) + 1.  With nshift = sigma - b + 16 and
// q = floor(fl(|hi| + |lo|) 2^nshift + 2^-30) + 1, |X'| < q 2^(b-16): the margin 2^-30 (in
// units of 2^(b-16)) covers the rounding of |hi| + |lo| (< 2^-37 there, since the scaled
// row maximum is below 2^b) and the integer rounding (< 2^(16-b)).  q <= 2^16 + 1,
// so the sum of q^2 over a row of up to 2^24 entries is an exact integer below 2^57, and
// ||X'_row||^2 < sum q^2 * 2^(2b-32).  Integer sums are independent of the order in which
// lanes and atomics add them, so the modulus count chosen from them is reproducible.
DDI unsigned long long norm_q2(double hi, double lo, int nshift)
{
    const double y = ldexp(fabs(hi) + fabs(lo), nshift) + 0x1p-30;
    const unsigned long long q = (unsigned long long)(uint32_t)y + 1ull;
    return q * q;
}

// Symmetric residue of c0 + c1*2^35 + c2*2^70 modulo m: r = s - m*rint(s/m), where
// s = c2*p2 + c1*p1 + c0 is an exact integer with |s| < 2^45.  The computed s*inv_m is
// within |s/m|*2^-52 < 2^-7/m of s/m.  For odd m, s/m is at least 1/(2m) away from every
// half-integer, so rint picks the same integer as for the exact quotient and r lies in
// [-(m-1)/2, (m-1)/2].  For m = 256 the product is exact and a tie gives r = +-128, whose
// byte 0x80 is -128 == 128 (mod 256).  Three FP64 operations and a conversion per entry
// and modulus; no integer fix-up.
DDI int sym_residue(double c0, double c1, double c2, const ModConst &mc)
{
    const double s = __builtin_fma(c2, mc.p2, __builtin_fma(c1, mc.p1, c0));
    const double q = rint(s * mc.inv_m);
    return (int)__builtin_fma(-q, mc.m, s);
}

// The low bytes of v0..v3 as one word, v0 in byte 0 (v_perm_b32: selector bytes 0-3 take
// bytes of the second source, 4-7 of the first, 0x0c gives zero).
DDI uint32_t pack_low_bytes(int v0, int v1, int v2, int v3)
{
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t)v1, (uint32_t)v0, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t)v3, (uint32_t)v2, 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Residue split of one lane's unit of a real operand: row r, k in [16*kg, 16*kg+16), all
// moduli, out[t*mod_stride + r*kp + k].  KC: entry (r,k) at X[r*ld + k] (row-contiguous);
// otherwise at X[k*ld + r].  NORM: returns the unit's part of the row's norm bound (the sum
// of norm_q2 over its entries k < K; 0 for a row holding a NaN or an infinity, whose
// integers are all zero).
template <bool KC, bool NORM = false>
DDI unsigned long long oz_split_unit(const ddv *__restrict__ X, int64_t ld, int64_t K,
                                     const unsigned long long *__restrict__ rmax, int bits,
                                     const ModConst *__restrict__ mcs, int nmod,
                                     int8_t *__restrict__ out, int64_t kp, int64_t mod_stride,
                                     int64_t r, int64_t kg)
{
    bool bad;
    const int sigma = row_sigma(rmax[r], bits, bad);
    double c0[16], c1[16], c2[16];
    unsigned long long nrm = 0ull;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int64_t k = kg * 16 + s;
        double hi = 0.0, lo = 0.0;
        if (k < K && !bad) {
            const ddv x = KC ? X[r * ld + k] : X[k * ld + r];
            hi = x.hi;
            lo = x.lo;
            if (NORM) nrm += norm_q2(hi, lo, sigma - bits + 16);
        }
        int_chunks(hi, lo, sigma, c0[s], c1[s], c2[s]);
    }
    for (int t = 0; t < nmod; ++t) {
        const ModConst mc = mcs[t];
        int res[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) res[s] = sym_residue(c0[s], c1[s], c2[s], mc);
        v4i w;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            w[q] = (int)pack_low_bytes(res[4 * q], res[4 * q + 1], res[4 * q + 2], res[4 * q + 3]);
        // non-temporal: the residue planes are read only after the whole split (by the residue
        // GEMMs), so they gain nothing from staying in L2 (the two 8192^3 splits: 1.395 ->
        // 1.35 ms)
        __builtin_nontemporal_store(w, reinterpret_cast<v4i *>(out + t * mod_stride + r * kp + kg * 16));
    }
    return nrm;
}

// Row maximum and residue split of a row-contiguous operand (entry (r,k) at X[r*ld + k]) in
// one pass over the row, for a group of L lanes per row (L a power of two dividing 64; the
// least with 8L >= Kc for short rows, see kc_lanes).  Lane ll of the group holds entries
// k0 + 8ll .. k0 + 8ll + 7 of each 8L-deep chunk: the group reads 16 * 8L contiguous bytes per
// chunk and writes 8L contiguous bytes per modulus, one 8-byte store per lane.  The row
// maximum (|hi| + |lo| as the bit pattern of a non-negative double, so a NaN or Inf
// dominates) is reduced across the group and stored by its first lane, the row's only writer
// (rmax needs no clearing); the residues then follow exactly as in oz_split_unit.  Rows of
// at most 8L entries stay in registers between the two passes; longer ones are read again.
// Entries in [K, Kc) get zero residues.  Every lane of the wave calls this (lanes of a row
// r >= R join the group shuffles only).  With rnorm, the first lane also stores the row's
// norm bound (the sum of norm_q2 over its entries k < K; 0 for a row holding a NaN or an
// infinity).
__host__ __device__ inline int kc_lanes(int64_t Kc)
{
    int L = 1;
    while (L < 64 && 8 * (int64_t)L < Kc) L *= 2;
    return L;
}

DDI void oz_rowsplit_kc_group(const ddv *__restrict__ X, int64_t ld, int64_t R, int64_t K,
                              int64_t Kc, int64_t r, int L, int ll,
                              unsigned long long *__restrict__ rmax, int bits,
                              const ModConst *__restrict__ mcs, int nmod,
                              int8_t *__restrict__ out, int64_t mod_stride,
                              unsigned long long *__restrict__ rnorm = nullptr)
{
    const bool live = r < R;
    const ddv *row = X + (live ? r : 0) * ld;
    constexpr int E = 8;
    const int64_t span = (int64_t)E * L;
    ddv v[E];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int64_t kk = k0 + E * ll + s;
            v[s] = live && kk < K ? row[kk] : ddv{0.0, 0.0};
        }
    };
    unsigned long long mx = 0ull;
    for (int64_t k0 = 0; k0 < K; k0 += span) {
        load(k0);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(mag(v[s]));
            mx = b > mx ? b : mx;
        }
    }
    for (int off = L / 2; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if (!live) return;
    if (ll == 0) rmax[r] = mx;
    bool bad;
    const int sigma = row_sigma(mx, bits, bad);
    int8_t *o = out + r * Kc;
    unsigned long long nrm = 0ull;
    for (int64_t k0 = 0; k0 < Kc; k0 += span) {
        if (K > span) load(k0);          // else the whole row is still in v
        const int64_t kl = k0 + E * ll;
        if (kl >= Kc) continue;
        double c0[E], c1[E], c2[E];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            int_chunks(bad ? 0.0 : v[s].hi, bad ? 0.0 : v[s].lo, sigma, c0[s], c1[s], c2[s]);
            if (rnorm && !bad && kl + s < K) nrm += norm_q2(v[s].hi, v[s].lo, sigma - bits + 16);
        }
        for (int t = 0; t < nmod; ++t) {
            const ModConst mc = mcs[t];
            int res[E];
#pragma unroll
            for (int s = 0; s < E; ++s) res[s] = sym_residue(c0[s], c1[s], c2[s], mc);
            v2i w;   // a non-temporal store, as in oz_split_unit
            w[0] = (int)pack_low_bytes(res[0], res[1], res[2], res[3]);
            w[1] = (int)pack_low_bytes(res[4], res[5], res[6], res[7]);
            __builtin_nontemporal_store(w, reinterpret_cast<v2i *>(o + t * mod_stride + kl));
        }
    }
    if (rnorm) {   // the row's norm bound (norm_q2), stored by the group's first lane
        for (int off = L / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, 64);
        if (ll == 0) rnorm[r] = nrm;
    }
}

// Complex residue split for the three-product (Karatsuba) form: row r of a complex operand,
// k in [8*kq, 8*kq+8) (8 entries per lane keep both parts' chunks in registers).  For each
// modulus t it writes three int8 planes at out + (3t+q)*plane + r*kp + k:
//   q = 0: re,  q = 1: sim*im,  q = 2: re + sim*im
// (sim = -1 conjugates).  The q = 2 residue is r0 + r1 folded once into the symmetric range
// (|r0 + r1| <= m - 1 for odd m; for m = 256 any representative's byte is the residue).
// Products of the planes of two operands give, modulo m, P0 = Re.Re', P1 = Im.Im' and
// P2 = (Re + Im)(Re' + Im'); the real and imaginary parts of the complex product are P0 - P1
// and P2 - P0 - P1.  KC: entry (r,k) at X[r*ld + k]; otherwise at X[k*ld + r].
template <bool KC>
DDI void oz_split_unit_k3(const zddv *__restrict__ X, int64_t ld, int64_t K,
                          const unsigned long long *__restrict__ rmax, int bits, int sim,
                          const ModConst *__restrict__ mcs, int nmod, int8_t *__restrict__ out,
                          int64_t kp, int64_t plane, int64_t r, int64_t kq)
{
    bool bad;
    const int sigma = row_sigma(rmax[r], bits, bad);
    double a0[8], a1[8], a2[8], b0[8], b1[8], b2[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int64_t k = kq * 8 + s;
        zddv v{{0.0, 0.0}, {0.0, 0.0}};
        if (k < K && !bad) v = KC ? X[r * ld + k] : X[k * ld + r];
        int_chunks(v.re.hi, v.re.lo, sigma, a0[s], a1[s], a2[s]);
        int_chunks(v.im.hi, v.im.lo, sigma, b0[s], b1[s], b2[s]);
        if (sim < 0) {
            b0[s] = -b0[s];
            b1[s] = -b1[s];
            b2[s] = -b2[s];
        }
    }
    const int64_t off = r * kp + kq * 8;
    for (int t = 0; t < nmod; ++t) {
        const ModConst mc = mcs[t];
        const int mi = (int)mc.m, h = (mi - 1) >> 1;
        int r0[8], r1[8], r2[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            r0[s] = sym_residue(a0[s], a1[s], a2[s], mc);
            r1[s] = sym_residue(b0[s], b1[s], b2[s], mc);
            const int x = r0[s] + r1[s];
            r2[s] = x > h ? x - mi : (x < -h ? x + mi : x);
        }
        int8_t *o = out + (int64_t)(3 * t) * plane + off;
        *reinterpret_cast<uint2 *>(o) =
            make_uint2(pack_low_bytes(r0[0], r0[1], r0[2], r0[3]), pack_low_bytes(r0[4], r0[5], r0[6], r0[7]));
        *reinterpret_cast<uint2 *>(o + plane) =
            make_uint2(pack_low_bytes(r1[0], r1[1], r1[2], r1[3]), pack_low_bytes(r1[4], r1[5], r1[6], r1[7]));
        *reinterpret_cast<uint2 *>(o + 2 * plane) =
            make_uint2(pack_low_bytes(r2[0], r2[1], r2[2], r2[3]), pack_low_bytes(r2[4], r2[5], r2[6], r2[7]));
    }
}

// The symmetric residue u = v - m*q, q = rint(v/m), of an exact INT32 dot product v
// (|v| <= 2^29: the residue GEMMs fold their sums below that), stored by the residue GEMMs
// as a signed byte.  The FMA t = v*fl(1/m) + 1.5*2^52 rounds the exact product v*fl(1/m)
// once, to the integer grid of [2^52, 2^53): v*fl(1/m) is within 2^-23 of v/m, never that
// close to a half-integer for odd m and exact for m = 256 (where a tie gives +-128, byte
// 0x80 either way), so q is the nearest integer to v/m; the low 32 bits of t hold q
// (2^51 + q, reduced mod 2^32) in two's complement.  |q| < 2^23, so u = v - q*m is one
// exact 24-bit multiply and a subtraction.  Four operations per entry, against six for
// floor(d*RU(1/m)) on an offset v with FP64 conversions both ways.  The reconstruction
// takes any representative of v's class modulo m (oz_crt_accumulate), so the canonical
// [0, m) form is not needed.  (The CRT weight y = (M/m)^-1 mod m lives in the
// reconstruction's fixed-point limbs.)
DDI int crt_residue(int v, const ModConst &mc)
{
    const double t = __builtin_fma((double)v, mc.inv_m, 0x1.8p52);
    return v - __mul24(__double2loint(t), (int)mc.m);
}

// The byte that the residue GEMMs store for v: crt_residue(v, mc) modulo 2^8, in three
// operations.  With neg_m = 2^24 - m (residue_neg_m), (q mod 2^24) * neg_m = -q*m modulo
// 2^24, so v + that product (one v_mad_u32_u24, which takes the low 24 bits of q by itself)
// agrees with u = v - q*m in its low 8 bits; the bits above them are not kept.
DDI uint32_t residue_neg_m(const ModConst &mc) { return 0x1000000u - (uint32_t)mc.m; }
DDI uint32_t crt_residue_byte(int v, const ModConst &mc, uint32_t neg_m)
{
    const double t = __builtin_fma((double)v, mc.inv_m, 0x1.8p52);
    return (uint32_t)v + __umul24((uint32_t)__double2loint(t), neg_m);
}
// The stored bytes of four entries as one word, v0 in byte 0.
DDI uint32_t residue_word(int v0, int v1, int v2, int v3, const ModConst &mc, uint32_t neg_m)
{
    return pack_low_bytes((int)crt_residue_byte(v0, mc, neg_m), (int)crt_residue_byte(v1, mc, neg_m),
                          (int)crt_residue_byte(v2, mc, neg_m), (int)crt_residue_byte(v3, mc, neg_m));
}

// A representative of v modulo m with |r| < 4m, from 32-bit arithmetic only (light on
// registers, for use inside the residue GEMMs' main loops).  For |v| <= 2^30 and m >= 71,
// v/m < 2^24; the three roundings of fl(fl(v) * fl(1/m)) (relative 2^-24 each) move it by
// less than 3 and the truncation by less than 1, so q is within 4 of v/m, and v - q*m is
// exact in INT32.  The class modulo m is unchanged, so the epilogue's crt_residue still
// returns the canonical residue of the whole sum.
DDI int fold_residue(int v, float inv_m, int m)
{
    const int q = (int)((float)v * inv_m);
    return v - q * m;
}

// Layout of the reduced products U (one uint8 per entry and product, rows and columns padded
// to multiples of 32, ldu = padded rows): the 32 x 32 blocks of an MFMA output tile, column
// block by column block, each block 1 KB in the order the v_mfma_i32_32x32x32_i8 accumulator
// holds it.  Word g*64 + l of a block (g = 0..3, lane l = 0..63) packs rows 8g + 4(l >> 5)
// .. +3 of column l & 31, so a wave stores 256 contiguous bytes per instruction (a
// column-major U took 32 separate 8-byte pieces per instruction, ~0.65 ms of the 8192^3 call).
// Byte offset of the word holding rows i .. i+3 (i a multiple of 4) of column j:
DDI int64_t u_word_offset(int64_t i, int64_t j, int64_t ldu)
{
    const int r = (int)(i & 31);
    return ((j >> 5) * (ldu >> 5) + (i >> 5)) * 1024 + (r >> 3) * 256 +
           ((j & 31) + 32 * ((r >> 2) & 1)) * 4;
}
// Byte offset of the 32 x 32 block whose first row and column are i0, j0 (multiples of 32).
DDI int64_t u_block_offset(int64_t i0, int64_t j0, int64_t ldu)
{
    return ((j0 >> 5) * (ldu >> 5) + (i0 >> 5)) * 1024;
}

// Workgroup -> output tile.  Workgroups are dispatched round-robin over the 8 XCDs (each
// with its own L2), so ids are first regrouped so that one XCD receives a contiguous run;
// runs are then walked in bands of GROUP tile-rows, column by column, so the workgroups
// resident on one XCD share A row-slabs and B column-slabs in its L2.
DDI void tile_order(int bid, int tiles_m, int tiles_n, int &tm, int &tn)
{
    constexpr int XCD = 8, GROUP = 4;
    const int T = tiles_m * tiles_n;
    int id = bid;
    if (T % XCD == 0) id = (bid % XCD) * (T / XCD) + bid / XCD;
    const int band = GROUP * tiles_n;
    const int first_m = (id / band) * GROUP;
    const int gm = (tiles_m - first_m < GROUP) ? tiles_m - first_m : GROUP;
    const int r = id % band;
    tm = first_m + r % gm;
    tn = r / gm;
}

// One BM x BN tile of nmod residue GEMMs: for t < nmod, U_t = (A_t B_t^T) mod m_t, A_t = Ab + t*sA: [Mp][kp] int8, B_t = Bb + t*sB: [Np][kp] int8, U_t = Ub + t*sU
// column-major [Np][ldu] uint8.  Four wave64 per workgroup in a 2 x 2 arrangement, each wave
// owns TM x TN MFMA tiles of 32 x 32 (v_mfma_i32_32x32x32_i8); 64-deep K slabs
// double-buffered in LDS with 80-byte row pitch (conflict-free b128 reads).  The slabs of
// all nmod products form one sequence, so the pipeline runs straight across modulus
// boundaries: the next modulus's first slab is in flight while the previous modulus's
// accumulators are reduced and stored (short K, as in batches of small problems, then
// costs no pipeline refill per modulus).
template <int TM, int TN, int BK = 64, bool FOLD = true>
DDI void oz_gemm_tile(const int8_t *__restrict__ Ab, const int8_t *__restrict__ Bb,
                      uint8_t *__restrict__ Ub, int64_t kp, int64_t ldu,
                      const ModConst *__restrict__ mcs, int nmod, int64_t sA, int64_t sB,
                      int64_t sU, int64_t i0, int64_t j0, int8_t (&sa)[2][64 * TM * (BK + 16)],
                      int8_t (&sb)[2][64 * TN * (BK + 16)])
{
    // Row pitch BK+16 bytes: 16 consecutive rows land in 16 distinct 16-byte bank slots
    // ((BK+16)/16 is odd), so the MFMA fragment reads are conflict-free.
    constexpr int PITCH = BK + 16, CPR = BK / 16;     // 16-byte chunks per row
    constexpr int LA = TM * BK / 64, LB = TN * BK / 64;   // 16-byte loads per lane and slab
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;

    v4i ra[LA], rb[LB];
#define OZ_GLOAD(PA, PB, K0)                                                                  \
    do {                                                                                      \
        _Pragma("unroll") for (int q = 0; q < LA; ++q) {                                      \
            const int c = tid + 256 * q;                                                      \
            ra[q] = *reinterpret_cast<const v4i *>((PA) + (int64_t)(c / CPR) * kp + (K0) + (c % CPR) * 16); \
        }                                                                                     \
        _Pragma("unroll") for (int q = 0; q < LB; ++q) {                                      \
            const int c = tid + 256 * q;                                                      \
            rb[q] = *reinterpret_cast<const v4i *>((PB) + (int64_t)(c / CPR) * kp + (K0) + (c % CPR) * 16); \
        }                                                                                     \
    } while (0)
#define OZ_LSTORE(BUF)                                                                        \
    do {                                                                                      \
        _Pragma("unroll") for (int q = 0; q < LA; ++q) {                                      \
            const int c = tid + 256 * q;                                                      \
            *reinterpret_cast<v4i *>(&sa[BUF][(c / CPR) * PITCH + (c % CPR) * 16]) = ra[q];   \
        }                                                                                     \
        _Pragma("unroll") for (int q = 0; q < LB; ++q) {                                      \
            const int c = tid + 256 * q;                                                      \
            *reinterpret_cast<v4i *>(&sb[BUF][(c / CPR) * PITCH + (c % CPR) * 16]) = rb[q];   \
        }                                                                                     \
    } while (0)

    v16i acc[TM][TN];
    auto clear = [&]() {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                acc[a][b] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    };
    // Reduce the accumulators of modulus t and store them, four rows per 32-bit word.
    auto flush = [&](int t) {
        const ModConst mc = mcs[t];
        const uint32_t neg_m = residue_neg_m(mc);
        // Word g of accumulator tile (a, b): rows 8g + 4(lane >> 5) of 32 x 32 block (a, b) of
        // the wave's sub-tile, column lane & 31, i.e. byte 4 * lane + 256 g of that block.
        const int64_t cstride = (ldu >> 5) * 1024;
        uint8_t *U = Ub + (int64_t)t * sU +
                     u_block_offset(i0 + wm * 32 * TM, j0 + wn * 32 * TN, ldu) + lane * 4;
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const uint32_t word = residue_word(acc[a][b][4 * g], acc[a][b][4 * g + 1],
                                                       acc[a][b][4 * g + 2], acc[a][b][4 * g + 3], mc, neg_m);
                    *reinterpret_cast<uint32_t *>(U + b * cstride + a * 1024 + g * 256) = word;
                }
    };
    // Deep products: every OZ_FOLD_DEPTH of depth the accumulators are replaced by a small
    // representative of their class modulo m_t, which keeps the INT32 sums exact at any depth.
    auto fold = [&](int t) {
        const float rm = (float)mcs[t].inv_m;
        const int mi = (int)mcs[t].m;
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[a][b][e] = fold_residue(acc[a][b][e], rm, mi);
    };
    constexpr int64_t FOLD_SLABS = OZ_FOLD_DEPTH / BK;
    clear();

    const int64_t nk = kp / BK;
    const int64_t S = nk * nmod;
    // Load cursor (the next slab to fetch) and compute cursor, as running counters.
    const int8_t *la = Ab, *lb = Bb;
    int64_t lkt = 0;
    auto advance = [&]() {
        if (++lkt == nk) {
            lkt = 0;
            la += sA;
            lb += sB;
        }
    };
    OZ_GLOAD(la, lb, lkt * BK);
    advance();
    OZ_LSTORE(0);
    __syncthreads();
    int ct = 0;
    int64_t ckt = 0;
    for (int64_t s = 0; s < S; ++s) {
        const int cur = (int)(s & 1);
        const bool more = s + 1 < S;
        if (more) {
            OZ_GLOAD(la, lb, lkt * BK);
            advance();
        }
        if (ckt == 0 && s > 0) {   // first slab of modulus ct: modulus ct-1 is complete
            flush(ct - 1);
            clear();
        }
        // Fragments are double-buffered: the LDS reads of step ks+1 are issued before the
        // MFMAs of step ks (scheduling barriers keep that order), so LDS latency overlaps
        // the matrix-core work instead of stalling it.
        constexpr int KS = BK / 32;
        v4i af[2][TM], bf[2][TN];
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            const int row = wm * 32 * TM + a * 32 + (lane & 31);
            af[0][a] = *reinterpret_cast<const v4i *>(&sa[cur][row * PITCH + (lane >> 5) * 16]);
        }
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int col = wn * 32 * TN + b * 32 + (lane & 31);
            bf[0][b] = *reinterpret_cast<const v4i *>(&sb[cur][col * PITCH + (lane >> 5) * 16]);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int cb = ks & 1, nb = cb ^ 1;
            if (ks + 1 < KS) {
#pragma unroll
                for (int a = 0; a < TM; ++a) {
                    const int row = wm * 32 * TM + a * 32 + (lane & 31);
                    af[nb][a] = *reinterpret_cast<const v4i *>(&sa[cur][row * PITCH + (ks + 1) * 32 + (lane >> 5) * 16]);
                }
#pragma unroll
                for (int b = 0; b < TN; ++b) {
                    const int col = wn * 32 * TN + b * 32 + (lane & 31);
                    bf[nb][b] = *reinterpret_cast<const v4i *>(&sb[cur][col * PITCH + (ks + 1) * 32 + (lane >> 5) * 16]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[cb][a], bf[cb][b], acc[a][b], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (more) OZ_LSTORE(cur ^ 1);
        __syncthreads();
        if (++ckt == nk) {
            ckt = 0;
            ++ct;
        } else if (FOLD && ckt % FOLD_SLABS == 0) {
            fold(ct);
        }
    }

#undef OZ_GLOAD
#undef OZ_LSTORE
    flush(nmod - 1);
    __syncthreads();   // the LDS slabs may be refilled by the caller's next tile
}

// ---------------------------------------------------------------------------------------
// 256 x 256 residue tile fed by direct global->LDS loads (global_load_lds_dwordx4).
//
// A ring of RING_NST LDS stages, each one 64-deep K slab of the A panel and of the B panel
// (2 x 256 rows x 64 bytes = 32 KB), is filled by DMA that never passes through VGPRs, so
// three stages are in flight while the fourth is consumed.  Eight waves (two per SIMD, so
// one wave's barrier or LDS wait is covered by its partner's MFMAs) own 64 x 128 sub-tiles
// (2 x 4 MFMA tiles of 32 x 32, or 4 x 8 of 16 x 16).  Within a stage row r, 16-byte chunk
// c is stored in a slot c ^ swz(r) chosen so that the fragment loads of the MFMA shape in use
// are free of bank conflicts, while each DMA instruction still writes 1 KB of contiguous LDS
// (16 rows x 4 chunks; the swizzle only permutes the global addresses its lanes read).
//
// Synchronisation: a wave's own DMA writes have landed when its vector-memory counter says
// so; the s_barrier then publishes them to the workgroup.  The barrier is bracketed by
// compiler-only memory clobbers instead of a fence, because a workgroup fence would also
// wait for the stages still in flight.
constexpr int RING_NST = 4, RING_BK = 64;
// LDS bytes of one ring stage (the A and the B slab) for BT x BT tiles.
template <int BT> constexpr int ring_slot() { return 2 * BT * RING_BK; }
constexpr int RING_SLOT = ring_slot<256>();
constexpr int RING_THREADS = 512;

// Wait until at most N of this wave's vector-memory operations are outstanding
// (gfx9 s_waitcnt encoding: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[5:4] at [15:14]).
template <int N> DDI void wait_vmem()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}
// Wait until at most n stages of DMA (Q instructions each) issued after the awaited one
// are outstanding.
template <int Q> DDI void wait_stages(int n)
{
    switch (n) {
    case 0: wait_vmem<0>(); break;
    case 1: wait_vmem<Q>(); break;
    case 2: wait_vmem<2 * Q>(); break;
    default: wait_vmem<3 * Q>(); break;
    }
}
DDI void wait_lds() { __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (0 << 8) | (3 << 14)); }

// The ring tile for a WM x WN grid of waves, each owning TM x TN MFMA tiles of 32 x 32: a
// BT x BT tile (BT = 32*TM*WM = 32*TN*WN), 64*WM*WN threads.  <4,2,2,4> is the 256 x 256
// tile of the single products (one workgroup per CU); <2,2,2,2> the 128 x 128 tile of the
// batched products (two per CU).  With MULTI, the tile runs the products of nmod moduli
// (A_t = Ab + t*sA, B_t = Bb + t*sB, U_t = Ub + t*sU, constants mcs[t]) as one sequence of
// stages, so the ring stays full across modulus boundaries; the accumulators of a modulus
// are reduced and stored while later stages' DMA is in flight.  With CLAMP, only the first
// rows_a rows of the A panel and rows_b rows of the B panel exist (multiples of 32): rows
// beyond them are fetched as copies of the last row (valid, cached memory) and their
// results are not stored.  With FOLD, the accumulators are reduced every OZ_FOLD_DEPTH of
// depth (needed only when kp exceeds it).  epi: when given, it replaces the store of each
// modulus's reduced products: epi(t, acc) receives modulus t's INT32 accumulators (the
// wave's TM x TN MFMA tiles) and U is not written.
//
// M16: the wave's 32TM x 32TN sub-tile runs on v_mfma_i32_16x16x64_i8 (2TM x 2TN tiles of
// 16 x 16, one 64-deep instruction each) instead of 32 x 32 x 32.  With random residues the
// 16 x 16 x 64 form sustains ~13% more operations per second under the chip's power limit
// (0.77 against 0.68 of the peak, register-only loops).  The LDS traffic per product is the
// same (it follows the sub-tile, not the instruction).  Step 0 of a stage multiplies the
// stage's A fragments by the first TN B fragments, step 1 by the other TN; the next stage's
// A fragments are read during step 1 into a second buffer.
//
// Walk (persistent workgroups, single products): instead of one tile, the workgroup runs the
// items walk.first, walk.first + walk.stride, ... (walk.count of them) of a launch's item
// space as one sequence of stages: item L is product L / tiles (modulus L / tiles / per)
// of tile L % tiles in tile_order, i.e. the order in which a grid of one workgroup per item
// would be dispatched, so the tiles resident on an XCD at a time are the same.  The ring is
// not drained and refilled between items and no workgroup is launched per item: the next
// item's first stages are in flight while the current item's accumulators are reduced and
// stored.  Ab, Bb, Ub, mcs are then the bases of product 0; i0, j0 and nmod are unused.
struct NoRingEpilogue {};
struct NoRingWalk {
    static constexpr int per = 1;
};
struct RingWalk {
    int first, stride, count;   // this workgroup's items
    int tiles_m, tiles, per;    // tiles per product (tiles_m of them per tile column)
    DDI void item(int q, int64_t kp, int64_t sA, int64_t sB, const int8_t *Ab, const int8_t *Bb,
                  const int8_t *&a, const int8_t *&b, int64_t &i0, int64_t &j0, int &t) const
    {
        const int L = first + q * stride;
        t = L / tiles;
        int tm, tn;
        tile_order(L - t * tiles, tiles_m, tiles / tiles_m, tm, tn);
        i0 = (int64_t)tm * 256;
        j0 = (int64_t)tn * 256;
        a = Ab + t * sA + i0 * kp;
        b = Bb + t * sB + j0 * kp;
    }
};
template <int WM, int WN, int TM, int TN, bool MULTI, bool CLAMP = false, bool FOLD = true,
          bool M16 = false, typename Epi = NoRingEpilogue, typename Walk = NoRingWalk,
          int NSTT = RING_NST>
DDI void oz_gemm_tile_ring(const int8_t *__restrict__ Ab, const int8_t *__restrict__ Bb,
                           uint8_t *__restrict__ Ub, int64_t kp, int64_t ldu,
                           const ModConst *__restrict__ mcs, int nmod, int64_t sA, int64_t sB,
                           int64_t sU, int64_t i0, int64_t j0, int8_t *smem, int rows_a = 0,
                           int rows_b = 0, Epi &&epi = Epi{}, const Walk &walk = Walk{})
{
    constexpr bool PERSIST = std::is_same<Walk, RingWalk>::value;
    static_assert(!PERSIST || (!MULTI && !CLAMP && 32 * TM * WM == 256), "persistent walk");
    constexpr int BT = 32 * TM * WM, NW = WM * WN;
    static_assert(BT == 32 * TN * WN, "square tile");
    constexpr int NST = NSTT, BK = RING_BK, SLOT = ring_slot<BT>();
    constexpr int DQ = BT / 16 / NW;            // DMA instructions per operand, wave and stage
    static_assert(DQ * NW * 16 == BT, "DMA rows");
    constexpr int NR = TM + TN, NM = TM * TN;   // fragment reads and MFMAs per 32-deep step
    static_assert(!M16 || std::is_same<std::decay_t<Epi>, NoRingEpilogue>::value,
                  "the epilogue receives 32 x 32 accumulator tiles");
    // w is uniform across the wave: taken through readfirstlane, the LDS addresses of the DMA
    // (one per wave and instruction) are scalar arithmetic.
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w % WM, wn = w / WM;
    const int64_t nk = kp / BK;
    int nitems = 1;
    if constexpr (PERSIST) nitems = walk.count;
    const int64_t S = MULTI ? nk * nmod : nk * nitems;   // stages of the whole sequence
    const int drow = lane >> 2, dslot = lane & 3;
    // Chunk c of stage row r sits in 16-byte slot c ^ swz(r).  Measured with
    // SQ_LDS_BANK_CONFLICT: (r >> 2) & 3 leaves the 32 x 32 operand reads (32 rows of one
    // chunk per 32 lanes) conflict-free, (r >> 1) & 3 the 16 x 16 x 64 ones (16 rows of four
    // chunks), which conflict 52M times per 13M reads with the other.
    auto swz = [](int r) { return M16 ? (r >> 1) & 3 : (r >> 2) & 3; };

    // DMA of the next stage into its ring slot (stages are issued in order, so the source
    // is a running cursor: modulus base pointers and the K slab within the modulus).
    // Instruction q of wave w covers rows (DQ*w + q) * 16 + lane / 4 of both panels.
    const int8_t *ca = Ab, *cb = Bb;
    int64_t ck = 0;
    // PERSIST: the output tile and product of the item being finished, and the items of the
    // DMA and compute cursors
    int64_t fi0 = i0, fj0 = j0;
    int ft = 0, dq = 0, cq = 0;
    if constexpr (PERSIST) walk.item(0, kp, sA, sB, Ab, Bb, ca, cb, fi0, fj0, ft);
    // Per-lane parts of the DMA source addresses (row and chunk within the panel), fixed for
    // the whole sequence; a stage adds the uniform panel base and K offset to them.
    int64_t loff_a[DQ], loff_b[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
        const int r = (DQ * w + q) * 16 + drow;
        const int c = dslot ^ swz(r);
        const int ra = CLAMP && r >= rows_a ? rows_a - 1 : r;
        const int rb = CLAMP && r >= rows_b ? rows_b - 1 : r;
        loff_a[q] = (int64_t)ra * kp + c * 16;
        loff_b[q] = (int64_t)rb * kp + c * 16;
    }
    auto issue = [&](int64_t s) {
        int8_t *slot = smem + (int)((uint32_t)s % NST) * SLOT;
        const int8_t *sa = ca + ck * BK, *sb = cb + ck * BK;
#pragma unroll
        for (int q = 0; q < DQ; ++q) {
            const int blk = DQ * w + q;
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(sa + loff_a[q]),
                (__attribute__((address_space(3))) void *)(slot + blk * 1024), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(sb + loff_b[q]),
                (__attribute__((address_space(3))) void *)(slot + BT * BK + blk * 1024), 16, 0, 0);
        }
        if (++ck == nk) {
            if constexpr (MULTI) {
                ck = 0;
                ca += sA;
                cb += sB;
            } else if constexpr (PERSIST) {
                ck = 0;
                if (++dq < nitems) {
                    int tt;
                    int64_t d0, d1;
                    walk.item(dq, kp, sA, sB, Ab, Bb, ca, cb, d0, d1, tt);
                }
            }
        }
    };
    v4i af[2][TM], bf[2][TN];
    // M16 fragments: a16 / bh = A and the last TN B fragments (16 columns each) of the stage
    // being multiplied, an16 / bl / bhn = A, the first and the last TN B fragments of the
    // next stage (bl is read only after step 0 has used it).
    v4i a16[2 * TM], an16[2 * TM], bl[TN], bh[TN], bhn[TN];
    // A 16-row fragment of a 64-deep stage: lane l holds row r0 + (l & 15), chunk l >> 4.
    auto frag16 = [&](const int8_t *base, int r0) {
        const int r = r0 + (lane & 15), c = lane >> 4;
        return *reinterpret_cast<const v4i *>(base + r * 64 + ((c ^ swz(r)) << 4));
    };
    // Fragments of 32-deep step ks (0 or 1) of stage s into fragment buffer d.  M16: all of
    // stage s's fragments (an16, bl, bhn), whatever ks and d.
    auto frag = [&](int64_t s, int ks, int d) {
        const int8_t *sl = smem + (int)((uint32_t)s % NST) * SLOT;
        if constexpr (M16) {
            (void)ks;
            (void)d;
#pragma unroll
            for (int a = 0; a < 2 * TM; ++a) an16[a] = frag16(sl, wm * 32 * TM + a * 16);
#pragma unroll
            for (int b = 0; b < TN; ++b) bl[b] = frag16(sl + BT * BK, wn * 32 * TN + b * 16);
#pragma unroll
            for (int b = 0; b < TN; ++b) bhn[b] = frag16(sl + BT * BK, wn * 32 * TN + (TN + b) * 16);
            return;
        }
        const int c = 2 * ks + (lane >> 5);
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            const int r = wm * 32 * TM + a * 32 + (lane & 31);
            af[d][a] = *reinterpret_cast<const v4i *>(sl + r * 64 + ((c ^ ((r >> 2) & 3)) << 4));
        }
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int r = wn * 32 * TN + b * 32 + (lane & 31);
            bf[d][b] = *reinterpret_cast<const v4i *>(sl + BT * BK + r * 64 + ((c ^ ((r >> 2) & 3)) << 4));
        }
    };
    // Accumulators: TM x TN tiles of 32 x 32 (16 per lane), or M16: 2TM x 2TN of 16 x 16 (4).
    using AccT = typename std::conditional<M16, v4i, v16i>::type;
    constexpr int AM = M16 ? 2 * TM : TM, AN = M16 ? 2 * TN : TN, AE = M16 ? 4 : 16;
    AccT acc[AM][AN];
    auto clear = [&]() {
#pragma unroll
        for (int a = 0; a < AM; ++a)
#pragma unroll
            for (int b = 0; b < AN; ++b)
#pragma unroll
                for (int e = 0; e < AE; ++e) acc[a][b][e] = 0;
    };
    // Z: the first stage of a product, whose MFMAs start from zero instead of the accumulators
    // (M16: both steps, which write disjoint accumulators; else step 0), so the accumulators
    // need no clearing after a flush.  Only the 32 x 32 x 32 tiles (the grouped products,
    // one product per modulus of a short depth) start that way: for the persistent walk,
    // whose items are 128 stages deep, the second copy of the stage loop cost 11% of the
    // 8192^3 residue GEMMs, against the 0.8% that the clearing takes.
    constexpr bool ZSTART = !PERSIST && !M16;
    auto mma = [&](int d, auto Z) {
        constexpr bool z = decltype(Z)::value;
        if constexpr (M16) {
            if (d == 0) {
#pragma unroll
                for (int a = 0; a < 2 * TM; ++a) a16[a] = an16[a];
#pragma unroll
                for (int b = 0; b < TN; ++b) bh[b] = bhn[b];
            }
            // Consecutive MFMAs share their B operand (measured 0.3-0.5% faster than sharing A).
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int a = 0; a < 2 * TM; ++a) {
                    AccT &c = acc[a][d * TN + b];
                    c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a16[a], d ? bh[b] : bl[b],
                                                             z ? AccT{0, 0, 0, 0} : c, 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                        af[d][a], bf[d][b], (z && d == 0) ? AccT{} : acc[a][b], 0, 0, 0);
        }
    };
    // Scheduling groups: each MFMA followed by one DMA (when refilling) and one of the step's
    // nr fragment reads.
    auto interleave = [&](bool with_dma, int nr) {
        constexpr int NMS = M16 ? 2 * TM * TN : NM;
#pragma unroll
        for (int i = 0; i < NMS; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            if (with_dma && i < 2 * DQ) __builtin_amdgcn_sched_group_barrier(0x010, 1, 0);
            if (i < nr) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
    };
    // Stage barrier: this wave's DMA of the awaited stage has landed (at most `later` newer
    // stages outstanding) and its fragment reads are complete; then all waves meet.
    auto stage_barrier = [&](int later) {
        wait_stages<2 * DQ>(later);
        wait_lds();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };
    // Reduce the accumulators of modulus t and store them, four rows per 32-bit word (or
    // hand them to the caller's epilogue).
    auto flush = [&](int t) {
        if constexpr (!std::is_same<std::decay_t<Epi>, NoRingEpilogue>::value) {
            epi(t, acc);
            return;
        }
        const ModConst mc = mcs[PERSIST ? t / walk.per : t];
        const uint32_t neg_m = residue_neg_m(mc);
        // Every word of the wave's sub-tile lies at a fixed offset from one per-lane base:
        // the wave's first 32 x 32 block, the lane's place in a block, then whole blocks
        // (1 KB down a block column, cstride bytes across block columns).  u_word_offset per
        // word cost ~10 instructions of 64-bit index arithmetic, more than the reduction.
        const int64_t cstride = (ldu >> 5) * 1024;
        const int lane_off = M16 ? (lane >> 5) * 256 + ((lane >> 4) & 1) * 128 + (lane & 15) * 4
                                 : lane * 4;
        uint8_t *U = Ub + (int64_t)t * sU +
                     u_block_offset(fi0 + wm * 32 * TM, fj0 + wn * 32 * TN, ldu) + lane_off;
        if constexpr (M16) {
            // Tile (a, b): lane l holds rows 4(l >> 4) .. +3 of column l & 15, i.e. rows
            // 16(a & 1) + 4(l >> 4) of 32 x 32 block (a >> 1, b >> 1), column 16(b & 1) + (l & 15).
#pragma unroll
            for (int a = 0; a < AM; ++a)
#pragma unroll
                for (int b = 0; b < AN; ++b) {
                    const int cl = wn * 32 * TN + b * 16 + (lane & 15);
                    const int rl = wm * 32 * TM + a * 16 + 4 * (lane >> 4);
                    if (CLAMP && (rl >= rows_a || cl >= rows_b)) continue;
                    const uint32_t word = residue_word(acc[a][b][0], acc[a][b][1], acc[a][b][2],
                                                       acc[a][b][3], mc, neg_m);
                    // non-temporal: U is read only after the launch (by the reconstruction),
                    // so its lines leave L2 to the residue panels
                    __builtin_nontemporal_store(word, reinterpret_cast<uint32_t *>(
                        U + (b >> 1) * cstride + (a >> 1) * 1024 + (a & 1) * 512 + (b & 1) * 64));
                }
        } else {
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) {
                    const int cl = wn * 32 * TN + b * 32 + (lane & 31);
                    const int rl = wm * 32 * TM + a * 32 + 4 * (lane >> 5);
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        // rows rl+8g .. +3 lie all inside or all outside rows_a (a multiple of 32)
                        if (CLAMP && (rl + 8 * g >= rows_a || cl >= rows_b)) continue;
                        const uint32_t word = residue_word(acc[a][b][4 * g], acc[a][b][4 * g + 1],
                                                           acc[a][b][4 * g + 2], acc[a][b][4 * g + 3], mc, neg_m);
                        __builtin_nontemporal_store(
                            word, reinterpret_cast<uint32_t *>(U + b * cstride + a * 1024 + g * 256));
                    }
                }
        }
    };
    // Deep products: every OZ_FOLD_DEPTH of depth the accumulators are replaced by a small
    // representative of their class modulo m_t, which keeps the INT32 sums exact at any depth.
    auto fold = [&](int t) {
        const int tc = PERSIST ? t / walk.per : t;
        const float rm = (float)mcs[tc].inv_m;
        const int mi = (int)mcs[tc].m;
#pragma unroll
        for (int a = 0; a < AM; ++a)
#pragma unroll
            for (int b = 0; b < AN; ++b)
#pragma unroll
                for (int e = 0; e < AE; ++e) acc[a][b][e] = fold_residue(acc[a][b][e], rm, mi);
    };
    constexpr int64_t FOLD_STAGES = OZ_FOLD_DEPTH / BK;
    // Compute cursor: modulus and K slab of the stage being finished.  The last stage of the
    // sequence is finished after the loops (flush of the last modulus).
    int ct = 0;
    int64_t ckt = 0;
    // Returns whether the accumulators were flushed (the next stage starts a product).
    auto stage_done = [&]() {
        ++ckt;
        if constexpr (MULTI) {
            if (ckt == nk) {   // the last stage of a modulus before the last one
                flush(ct);
                if constexpr (!ZSTART) clear();
                ckt = 0;
                ++ct;
                return true;
            }
        }
        if constexpr (PERSIST) {
            if (ckt == nk) {   // the last stage of an item before the last one
                flush(ft);
                if constexpr (!ZSTART) clear();
                ckt = 0;
                const int8_t *a, *b;
                walk.item(++cq, kp, sA, sB, Ab, Bb, a, b, fi0, fj0, ft);
                return true;
            }
        }
        if (FOLD && ckt % FOLD_STAGES == 0) fold(PERSIST ? ft : ct);
        return false;
    };
    using ZERO = std::true_type;
    using ACC = std::false_type;
    bool fresh = true;   // the next stage starts a product
    if constexpr (!ZSTART) clear();

    const int64_t npro = S < NST ? S : NST;
    for (int64_t s = 0; s < npro; ++s) issue(s);
    stage_barrier((int)(npro - 1));
    int64_t kt = 0;
    if constexpr (M16) {
        // All of stage kt + 1's fragments are read during step 1 of stage kt, after the
        // barrier that published it, into buffers that step 1 does not read (the last B half
        // has its own: bhn, copied to bh at the next step 0).  So the reads issue early among
        // step 1's MFMAs, and the next barrier, two MFMAs into step 0, finds them complete.
        // Reading the last B half at the start of step 0 instead left every wave waiting out
        // the LDS latency in front of that barrier (the kept form measured 1.7-2.2% faster on
        // three MI355X GPUs).
        frag(0, 0, 0);
        // Stage pairs in ping-pong buffers: stage kt's A and last-B fragments in pa[kt & 1],
        // ph[kt & 1], so no fragment is copied between stages (the single-buffer form below
        // moves 16 register pairs per stage; the pairs measured 1.6-3.0% faster on two
        // MI355X GPUs).  The remaining stages (refills left over from the pairs, then the drain)
        // take the single-buffer form, handed over through an16 / bhn.
        {
            using I0 = std::integral_constant<int, 0>;
            using I1 = std::integral_constant<int, 1>;
            v4i pa[2][2 * TM], ph[2][TN];
#pragma unroll
            for (int a = 0; a < 2 * TM; ++a) pa[0][a] = an16[a];
#pragma unroll
            for (int b = 0; b < TN; ++b) ph[0][b] = bhn[b];
            auto pmma = [&](auto P, int d) {
                constexpr int p = decltype(P)::value;
#pragma unroll
                for (int b = 0; b < TN; ++b)
#pragma unroll
                    for (int a = 0; a < 2 * TM; ++a) {
                        AccT &c = acc[a][d * TN + b];
                        c = __builtin_amdgcn_mfma_i32_16x16x64_i8(pa[p][a], d ? ph[p][b] : bl[b], c, 0, 0, 0);
                    }
            };
            auto pfrag = [&](auto P, int64_t st) {
                constexpr int p = decltype(P)::value;
                const int8_t *sl = smem + (int)((uint32_t)st % NST) * SLOT;
#pragma unroll
                for (int a = 0; a < 2 * TM; ++a) pa[p][a] = frag16(sl, wm * 32 * TM + a * 16);
#pragma unroll
                for (int b = 0; b < TN; ++b) bl[b] = frag16(sl + BT * BK, wn * 32 * TN + b * 16);
#pragma unroll
                for (int b = 0; b < TN; ++b) ph[p][b] = frag16(sl + BT * BK, wn * 32 * TN + (TN + b) * 16);
            };
            auto pstage = [&](auto P, int64_t k) {
                constexpr int p = decltype(P)::value;
                pmma(P, 0);
                interleave(false, 0);
                stage_barrier(NST - 2);
                issue(k + NST);
                pfrag(std::integral_constant<int, 1 - p>{}, k + 1);
                pmma(P, 1);
                interleave(true, 2 * TM + 2 * TN);
                stage_done();
            };
            for (; kt + 1 + NST < S; kt += 2) {
                pstage(I0{}, kt);
                pstage(I1{}, kt + 1);
            }
#pragma unroll
            for (int a = 0; a < 2 * TM; ++a) an16[a] = pa[0][a];
#pragma unroll
            for (int b = 0; b < TN; ++b) bhn[b] = ph[0][b];
        }
        for (; kt + NST < S; ++kt) {
            mma(0, ACC{});
            interleave(false, 0);
            stage_barrier(NST - 2);
            issue(kt + NST);
            frag(kt + 1, 0, 0);
            mma(1, ACC{});
            interleave(true, 2 * TM + 2 * TN);
            stage_done();
        }
        for (; kt + 1 < S; ++kt) {
            mma(0, ACC{});
            interleave(false, 0);
            const int64_t later = S - kt - 2;
            stage_barrier((int)(later < NST - 2 ? later : NST - 2));
            frag(kt + 1, 0, 0);
            mma(1, ACC{});
            interleave(false, 2 * TM + 2 * TN);
            stage_done();
        }
        mma(0, ACC{});
        interleave(false, 0);
        mma(1, ACC{});
    } else {
        frag(0, 0, 0);
        // Steady state.  Step 0 of stage kt runs while step 1's fragments are read; the barrier
        // before step 1 frees slot kt (its last fragments are in registers), which is refilled
        // with stage kt + NST while step 1 runs and stage kt + 1's first fragments are read.
        auto steady = [&](auto Z) {
            frag(kt, 1, 1);
            mma(0, Z);
            interleave(false, NR);
            stage_barrier(NST - 2);
            issue(kt + NST);
            frag(kt + 1, 0, 0);
            mma(1, Z);
            interleave(true, NR);
        };
        for (; kt + NST < S; ++kt) {
            if (ZSTART && fresh) steady(ZERO{});
            else steady(ACC{});
            fresh = stage_done();
        }
        // Drain: the last NST - 1 stage transitions issue no refills.
        auto drain = [&](auto Z) {
            frag(kt, 1, 1);
            mma(0, Z);
            interleave(false, NR);
            const int64_t later = S - kt - 2;
            stage_barrier((int)(later < NST - 2 ? later : NST - 2));
            frag(kt + 1, 0, 0);
            mma(1, Z);
            interleave(false, NR);
        };
        for (; kt + 1 < S; ++kt) {
            if (ZSTART && fresh) drain(ZERO{});
            else drain(ACC{});
            fresh = stage_done();
        }
        auto last = [&](auto Z) {
            frag(kt, 1, 1);
            mma(0, Z);
            interleave(false, NR);
            wait_lds();
            mma(1, Z);
        };
        if (ZSTART && fresh) last(ZERO{});
        else last(ACC{});
    }
    flush(MULTI ? nmod - 1 : (PERSIST ? ft : 0));
    __syncthreads();   // the ring may be refilled by the caller's next tile
}

// 128-bit fixed-point fraction sum_t u_t * round(2^128/m_t) mod 2^128, from four exact FP64
// limb accumulators (each < 2^45), as a signed double-double value F (|F| <= 2^127).
DDI ddv fixed128_to_dd(double a0, double a1, double a2, double a3)
{
    double c = floor(a0 * 0x1p-32);
    const double f0 = __builtin_fma(-c, 0x1p32, a0);
    a1 += c;
    c = floor(a1 * 0x1p-32);
    const double f1 = __builtin_fma(-c, 0x1p32, a1);
    a2 += c;
    c = floor(a2 * 0x1p-32);
    const double f2 = __builtin_fma(-c, 0x1p32, a2);
    a3 += c;
    c = floor(a3 * 0x1p-32);
    double f3 = __builtin_fma(-c, 0x1p32, a3);
    if (f3 >= 0x1p31) f3 -= 0x1p32;              // symmetric range: F - 2^128
    ddv v;
    dd_two_sum(f3 * 0x1p96, f2 * 0x1p64, v.hi, v.lo);
    v = dd_add_exact(v, ddv{f1 * 0x1p32, 0.0});
    v = dd_add_exact(v, ddv{f0, 0.0});
    return v;
}

DDI ddv reconstruct(const double acc[4], double Mhi, double Mlo, int e, bool bad)
{
    if (bad) return ddv{__builtin_nan(""), __builtin_nan("")};
    const ddv f = fixed128_to_dd(acc[0], acc[1], acc[2], acc[3]);
    const ddv c = dd_mul_exact(f, ddv{Mhi, Mlo});
    return ddv{ldexp(c.hi, e), ldexp(c.lo, e)};
}

// Reconstruction + alpha/beta update of rows i0..i0+3 (those < m) of column j.
// Fixed-point CRT sums of RPL consecutive rows i0.. of column j: acc[e][p][l] accumulates
// u_t * limb l of W_t = round(2^128 y_t / m_t) over the moduli (part p: real / imaginary).
// Any representative u_t of the class of C' modulo m_t will do: u_t + k m_t changes the sum
// by k m_t W_t = k y_t 2^128 + k m_t (W_t - 2^128 y_t / m_t), i.e. by at most |k| m_t / 2
// units of 2^-128 of M, so the residue GEMMs' signed bytes (|u_t| <= 128) and the complex
// combinations below (|u_t| <= 384) are used as they are; every limb product stays below
// 2^41 and every limb sum below 2^47, exact in FP64.
template <bool CPLX, int RPL>
DDI void oz_crt_accumulate(const uint8_t *__restrict__ U, int64_t ldu, int64_t sU, int nmod,
                           const ModConst *__restrict__ mcs, int64_t j, int64_t i0,
                           double (&acc)[RPL][CPLX ? 2 : 1][4])
{
    static_assert(RPL == 4, "rows per lane");
    constexpr int P = CPLX ? 2 : 1;
    // Complex: three product planes per modulus (P0, P1, P2 of oz_split_unit_k3), from which
    // the real part P0 - P1 and the imaginary part P2 - P0 - P1 are formed modulo m.
    constexpr int NQ = CPLX ? 3 : 1;
#pragma unroll
    for (int e = 0; e < RPL; ++e)
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int l = 0; l < 4; ++l) acc[e][p][l] = 0.0;
    // One modulus: the residue bytes of RPL rows times the four weight limbs.
    // byte e of a word as a signed integer (the residue GEMMs store symmetric residues)
    auto sbyte = [](uint32_t word, int e) { return ((int)(word << (24 - 8 * e))) >> 24; };
    auto step = [&](int t, const uint32_t (&word)[NQ]) {
        const double w0 = mcs[t].w[0], w1 = mcs[t].w[1], w2 = mcs[t].w[2], w3 = mcs[t].w[3];
#pragma unroll
        for (int e = 0; e < RPL; ++e) {
            int u[P];
            if constexpr (CPLX) {
                const int p0 = sbyte(word[0], e), p1 = sbyte(word[1], e), p2 = sbyte(word[2], e);
                u[0] = p0 - p1;               // representatives of Re and Im modulo m,
                u[P - 1] = p2 - p0 - p1;      // |u| <= 384
            } else {
                u[0] = sbyte(word[0], e);
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double d = (double)u[p];
                acc[e][p][0] = __builtin_fma(d, w0, acc[e][p][0]);
                acc[e][p][1] = __builtin_fma(d, w1, acc[e][p][1]);
                acc[e][p][2] = __builtin_fma(d, w2, acc[e][p][2]);
                acc[e][p][3] = __builtin_fma(d, w3, acc[e][p][3]);
            }
        }
    };
    auto load = [&](int t, uint32_t (&word)[NQ]) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            word[q] = *reinterpret_cast<const uint32_t *>(U + (int64_t)(t * NQ + q) * sU +
                                                          u_word_offset(i0, j, ldu));
    };
    // The residue words of TB moduli are loaded before any of their FMAs, so each lane
    // keeps TB*NQ loads in flight.  The limb sums are exact integers, so the grouping does
    // not change the result.
    constexpr int TB = 8;
    int t = 0;
    for (; t + TB <= nmod; t += TB) {
        uint32_t words[TB][NQ];
#pragma unroll
        for (int q = 0; q < TB; ++q) load(t + q, words[q]);
#pragma unroll
        for (int q = 0; q < TB; ++q) step(t + q, words[q]);
    }
    for (; t < nmod; ++t) {
        uint32_t word[NQ];
        load(t, word);
        step(t, word);
    }
}

// alpha * (A B^T)_ij from the CRT sums of entry (i, j), with the row/column scalings undone;
// NaN when row i of A or column j of B holds a NaN or an infinity.
DDI ddv oz_crt_value(const double (&acc)[1][4], ddv alpha, double Mhi, double Mlo, int ex, bool bad)
{
    return dd_mul_exact(alpha, reconstruct(acc[0], Mhi, Mlo, ex, bad));
}
DDI zddv oz_crt_value(const double (&acc)[2][4], zddv alpha, double Mhi, double Mlo, int ex, bool bad)
{
    const zddv cab{reconstruct(acc[0], Mhi, Mlo, ex, bad), reconstruct(acc[1], Mhi, Mlo, ex, bad)};
    return zdd_mul_exact(alpha, cab);
}

// C := beta*C + v for one entry; with beta == 0, C is not read (a NaN already in C must
// not propagate).
DDI void oz_crt_store(ddv *cp, ddv v, ddv beta, int beta_mode)
{
    ddv c0{0.0, 0.0};
    if (beta_mode != BETA_ZERO) c0 = dd_beta_prepass(*cp, beta, beta_mode);
    *cp = dd_add_exact(c0, v);
}
DDI void oz_crt_store(zddv *cp, zddv v, zddv beta, int beta_mode)
{
    zddv c0{{0.0, 0.0}, {0.0, 0.0}};
    if (beta_mode != BETA_ZERO) c0 = zdd_beta_prepass(*cp, beta, beta_mode);
    *cp = zdd_add_exact(c0, v);
}

// Reconstruction of one 32 x 32 block (bi, bj) of C by a workgroup of 256 lanes.  Wave w,
// lane l take rows 32bi + 8w + 4(l >> 5) .. +3 of column 32bj + (l & 31): their U words
// (u_word_offset) are 256 contiguous bytes per wave and product.  The values alpha * (A B^T)
// are staged in LDS (pitch 33 entries: the column-strided writes spread over the banks), then
// C is read (beta != 0) and written with consecutive lanes on consecutive rows, 32 rows of a
// column contiguous.  Every lane of the workgroup must call this (it holds a barrier).
template <bool CPLX>
DDI void oz_crt_block(const uint8_t *__restrict__ U, int64_t ldu, int64_t sU, int nmod,
                      const ModConst *__restrict__ mcs, double Mhi, double Mlo,
                      const unsigned long long *__restrict__ rmaxA,
                      const unsigned long long *__restrict__ rmaxB, int bits, int64_t m,
                      int64_t n, typename std::conditional<CPLX, zddv, ddv>::type alpha,
                      typename std::conditional<CPLX, zddv, ddv>::type beta, int beta_mode,
                      typename std::conditional<CPLX, zddv, ddv>::type *__restrict__ C,
                      int64_t ldc, int64_t bi, int64_t bj,
                      typename std::conditional<CPLX, zddv, ddv>::type (&stage)[32][33])
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t j = bj * 32 + (l & 31);
    const int r0 = 8 * w + 4 * (l >> 5);
    const int64_t i0 = bi * 32 + r0;
    if (i0 < m && j < n) {
        double acc[4][CPLX ? 2 : 1][4];
        oz_crt_accumulate<CPLX, 4>(U, ldu, sU, nmod, mcs, j, i0, acc);
        bool badj;
        const int tau = row_sigma(rmaxB[j], bits, badj);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t i = i0 + e < m ? i0 + e : m - 1;
            bool badi;
            const int sigma = row_sigma(rmaxA[i], bits, badi);
            stage[l & 31][r0 + e] =
                oz_crt_value(acc[e], alpha, Mhi, Mlo, -128 - sigma - tau, badi || badj);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int idx = threadIdx.x + 256 * it, c = idx >> 5, r = idx & 31;
        const int64_t i = bi * 32 + r, jj = bj * 32 + c;
        if (i < m && jj < n) oz_crt_store(&C[jj * ldc + i], stage[c][r], beta, beta_mode);
    }
}

}  // namespace mw
