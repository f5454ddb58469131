// dd_device.hpp -- double-double arithmetic on the CDNA4 FP64 vector ALU.
//
// Every operation here is an explicit IEEE binary64 op (v_add_f64 / v_mul_f64 /
// v_fma_f64, round-to-nearest-even, FP64 denormals preserved).  The translation unit
// must be compiled with -ffp-contract=off so that no a*b+c pair is fused behind our
// back; the FMAs the algorithms need are written with __builtin_fma.
//
// Two operation sets, matching the reference's numerics op for op:
//   fast      : DWTimesDW with FMA (7 ops) and SloppyDWPlusDW (11 ops), the GEMM inner
//               loop vocabulary (src/epblas-parallel/multifloats/mf_simd_fast.h:49-144);
//   faithful  : the documented sequences of the multifloats scalar operator* / operator+
//               (mf_simd_exact.h:60-99), used for the beta pre-pass and the C update.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

struct ddv { double hi, lo; };
struct zddv { ddv re, im; };

#define DDI __device__ __forceinline__

// TwoProd via FMA: p + e == a*b exactly.
DDI void dd_two_prod(double a, double b, double &p, double &e)
{
    p = a * b;
    e = __builtin_fma(a, b, -p);
}

// Knuth TwoSum (6 ops, no ordering precondition).
DDI void dd_two_sum(double a, double b, double &s, double &e)
{
    double t = a + b;
    double bv = t - a;
    double av = t - bv;
    e = (a - av) + (b - bv);
    s = t;
}

// Dekker Fast2Sum (3 ops, |a| >= |b|).
DDI void dd_fast_two_sum(double a, double b, double &s, double &e)
{
    double t = a + b;
    e = b - (t - a);
    s = t;
}

// fast a*b: cross terms fused into the TwoProd error, a.hi*b.lo first.
DDI ddv dd_mul_fast(ddv a, ddv b)
{
    double p, e;
    dd_two_prod(a.hi, b.hi, p, e);
    e = __builtin_fma(a.hi, b.lo, e);
    e = __builtin_fma(a.lo, b.hi, e);
    ddv r;
    dd_fast_two_sum(p, e, r.hi, r.lo);
    return r;
}

// fast a+b ("sloppy": one TwoSum on the high limbs, low limbs added plainly).
DDI ddv dd_add_fast(ddv a, ddv b)
{
    double s, e;
    dd_two_sum(a.hi, b.hi, s, e);
    e = e + (a.lo + b.lo);
    ddv r;
    dd_fast_two_sum(s, e, r.hi, r.lo);
    return r;
}

// faithful a*b: cross terms as separate products added without fusion.
DDI ddv dd_mul_exact(ddv a, ddv b)
{
    double p, e;
    dd_two_prod(a.hi, b.hi, p, e);
    double c1 = a.hi * b.lo;
    double c2 = a.lo * b.hi;
    e = e + (c1 + c2);
    ddv r;
    dd_fast_two_sum(p, e, r.hi, r.lo);
    return r;
}

// faithful a+b: TwoSum on both limb pairs, Fast2Sum combine, two error adds.
DDI ddv dd_add_exact(ddv a, ddv b)
{
    double s, e, t, f;
    dd_two_sum(a.hi, b.hi, s, e);
    dd_two_sum(a.lo, b.lo, t, f);
    dd_fast_two_sum(s, t, s, t);
    e = e + f;
    e = e + t;
    ddv r;
    dd_fast_two_sum(s, e, r.hi, r.lo);
    return r;
}

DDI ddv dd_neg(ddv a) { return ddv{-a.hi, -a.lo}; }
DDI ddv dd_sub_exact(ddv a, ddv b) { return dd_add_exact(a, dd_neg(b)); }

// Complex: fast product/sum for the inner loop, faithful product/sum for the C update.
DDI zddv zdd_mul_fast(zddv a, zddv b)
{
    zddv r;
    r.re = dd_add_fast(dd_mul_fast(a.re, b.re), dd_neg(dd_mul_fast(a.im, b.im)));
    r.im = dd_add_fast(dd_mul_fast(a.re, b.im), dd_mul_fast(a.im, b.re));
    return r;
}
DDI zddv zdd_add_fast(zddv a, zddv b)
{
    return zddv{dd_add_fast(a.re, b.re), dd_add_fast(a.im, b.im)};
}
DDI zddv zdd_mul_exact(zddv a, zddv b)
{
    zddv r;
    r.re = dd_sub_exact(dd_mul_exact(a.re, b.re), dd_mul_exact(a.im, b.im));
    r.im = dd_add_exact(dd_mul_exact(a.re, b.im), dd_mul_exact(a.im, b.re));
    return r;
}
DDI zddv zdd_add_exact(zddv a, zddv b)
{
    return zddv{dd_add_exact(a.re, b.re), dd_add_exact(a.im, b.im)};
}
DDI zddv zdd_conj(zddv a) { return zddv{a.re, dd_neg(a.im)}; }

// beta pre-pass modes (mgemm_parallel.cpp:68-77): 0 -> exact zero, 1 -> keep, 2 -> scale.
enum { BETA_ZERO = 0, BETA_ONE = 1, BETA_SCALE = 2 };

DDI ddv dd_beta_prepass(ddv c, ddv beta, int mode)
{
    if (mode == BETA_ZERO) return ddv{0.0, 0.0};
    if (mode == BETA_ONE) return c;
    return dd_mul_exact(c, beta);
}
DDI zddv zdd_beta_prepass(zddv c, zddv beta, int mode)
{
    if (mode == BETA_ZERO) return zddv{{0.0, 0.0}, {0.0, 0.0}};
    if (mode == BETA_ONE) return c;
    return zdd_mul_exact(c, beta);
}
