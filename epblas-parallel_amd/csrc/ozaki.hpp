// ozaki.hpp -- the exact-product double-double GEMM (Ozaki scheme II: integer residues on
// the INT8 matrix cores, Chinese-remainder reconstruction in FP64).
#pragma once
#include "mw_internal.hpp"

namespace mw {

int record_error(int code, const char *fmt, ...);

constexpr int OZ_MAX_MOD = 40;

// Per-modulus constants, read by every Ozaki kernel (uniform across a wave).
struct ModConst {
    double m, inv_m;      // modulus and 1/m (rounded)
    double y;             // ((M/m)^-1 mod m): CRT weight numerator
    double p1, p2;        // 2^35 mod m, 2^70 mod m: chunk weights for the residue split
    double w[4];          // round(2^128/m) as four 32-bit limbs (w[0] least significant)
};

// One reconstruction configuration: the first nmod moduli of the table, and the integer
// width b every scaled operand row/column is rounded to (|A'|, |B'| <= 2^b).
struct OzPlan {
    int nmod;
    int bits;
    double log2M;
    double Mhi, Mlo;      // product of the moduli as a double-double
    ModConst mc[OZ_MAX_MOD];
};

// Choose the plan for a contraction depth k (k_eff = k for real, 2k for complex): the
// fewest moduli whose product M satisfies M/2 > k_eff * 2^(2b) with b >= target bits.
// Returns false when k is out of range.
bool ozaki_make_plan(int64_t k_eff, int target_bits, int force_nmod, OzPlan *plan);

// Reconstruction constants of the first n moduli, n = 1..OZ_MAX_MOD, beside the set's table
// of ModConst (the CRT weights depend on the whole set): M as a double-double and log2 M.
// Entry n of a device array indexed by the modulus count.
struct OzSet {
    double Mhi, Mlo, log2M;
};

// Data-dependent modulus count (real products).  The split runs with a fixed integer width
// b and the worst-case count (the plan of ozaki_make_fixed_plan); the row norm bounds it
// accumulates then bound every |C'_ij| by Cauchy-Schwarz, and a device-side choice takes the
// fewest moduli whose product exceeds twice that bound.  The residue GEMMs of the moduli
// beyond it return at once, and the reconstruction uses the chosen set.  Mode 1 (default):
// for products of at least OZ_ESTIMATE_MIN_MNK multiply-adds, where a modulus saved (1/29
// of the residue GEMMs) outweighs the extra launch that chooses; 2: every real product;
// 0 (MWGEMM_OZAKI_ESTIMATE=0, or a forced modulus count): never.
constexpr double OZ_ESTIMATE_MIN_MNK = 8589934592.0;   // 2^33, e.g. 2048^3
int ozaki_estimate();
void ozaki_set_estimate(int mode);
// Whether a real product of m x n x k uses the data-dependent count.
bool ozaki_estimates(int64_t m, int64_t n, int64_t k);
// Integer width of the estimating plans: the configured target, else 102 (normwise error
// bound 2^-101 + the reconstruction's few 2^-104 < 1e-30).
int ozaki_estimate_bits();
// The plan with integer width exactly b and the fewest moduli for the worst case
// M/2 > k_eff * 2^(2b).
bool ozaki_make_fixed_plan(int64_t k_eff, int bits, OzPlan *plan);
// The modulus count chosen by the last estimating call on the current device (waits for
// the device); 0 when none has run.
int ozaki_last_estimate(int *nmod);
// Test hook: added to every data-dependent count (clamped to [1, worst case]), so a test can
// show that one modulus fewer than the chosen count breaks an adversarial product.
void ozaki_set_count_offset(int offset);
int ozaki_count_offset();

// Depth after which the residue GEMMs reduce their INT32 accumulators modulo m: a sum of
// OZ_FOLD_DEPTH products of symmetric residues (|r| <= 128) stays below 2^29, so a running
// residue in [0, m) plus the next partial sum stays inside the 2^30 range the epilogue's
// exact quotient handles.  Any depth is thus exact.
constexpr int64_t OZ_FOLD_DEPTH = 32768;
// Largest contraction depth (k for real, 2k for complex) the plans cover: the modulus count
// grows by one bit per doubling of k, and 2^24 still fits the 40-modulus table at b = 103.
constexpr int64_t OZ_MAX_KEFF = int64_t(1) << 24;

bool ozaki_supported(int64_t k);
bool ozaki_supported_cplx(int64_t k);

// Return value of ozaki_gemm_real / ozaki_gemm_cplx when the call's workspace (or modulus
// table) could not be allocated: nothing of the call was enqueued, and the caller runs the
// plain path instead.
constexpr int OZ_NO_WORKSPACE = 1;
// a_tag != 0: the caller runs a sequence of calls that share op(A) (the column panels of one
// host-pointer GEMM) and names it by a_tag.  A call whose tag matches the last call's on this
// device (real or complex alike), with the same m, k, modulus count and integer width and an
// unmoved workspace, reuses op(A)'s maxima and residues instead of recomputing them; any
// other call recomputes them.
int ozaki_gemm_real(const GemmArgsR &g, hipStream_t s, uint64_t a_tag = 0);
// The same product with its workspace in the library's static per-device scratch memory
// (NOALLOC_SCRATCH_BYTES, no device allocation).  A product whose workspace does not fit
// returns OZ_NO_WORKSPACE when may_fall_back, else an error.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
int ozaki_gemm_real_noalloc(const GemmArgsR &g, hipStream_t s, bool may_fall_back);
int ozaki_gemm_cplx(const GemmArgsZ &g, hipStream_t s, uint64_t a_tag = 0);

// Test hook: U[j*ldu + i] = (sum_k A[i*kp + k] * B[j*kp + k]) mod m for one
// modulus of the table, computed by the production INT8 MFMA kernel of the given tile
// variant (0..3).  mp and np must be multiples of 256, kp a multiple of 64.
int ozaki_residue_gemm_test(const int8_t *dA, const int8_t *dB, uint8_t *dU, int64_t mp,
                            int64_t np, int64_t kp, int mod_index, int variant, hipStream_t s);

// Current target bit width and a forced modulus count (0 = automatic).
int ozaki_target_bits();
int ozaki_forced_nmod();
void ozaki_set_params(int bits, int nmod);

}  // namespace mw
