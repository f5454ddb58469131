"""Host-side checks of the exact-product (Ozaki scheme II) path that need no GPU: the
modulus table, the CRT weights the library computes, the modulus count it picks, and the
fixed-point reconstruction identity its kernels evaluate (checked with Python integers)."""
import math
import random

import pytest

import mwblas


def test_moduli_pairwise_coprime_and_weights_are_inverses():
    m, y = mwblas.ozaki_moduli(40)
    m = [int(v) for v in m]
    for i in range(len(m)):
        assert 2 <= m[i] <= 256
        for j in range(i):
            assert math.gcd(m[i], m[j]) == 1
    for n in (8, 29, 40):
        mm, yy = mwblas.ozaki_moduli(n)
        M = math.prod(int(v) for v in mm)
        for mt, yt in zip(mm, yy):
            assert ((M // int(mt)) * int(yt)) % int(mt) == 1


@pytest.mark.parametrize("k,cplx", [(1, False), (512, False), (8192, False), (65536, False),
                                    (70000, False), (1 << 20, False), (4096, True),
                                    (32768, True), (40000, True)])
@pytest.mark.parametrize("estimate", [True, False])
def test_plan_covers_the_product_range(k, cplx, estimate):
    """The count covers the worst case |C'| <= k_eff 2^(2b) and one fewer would not.  Real
    products with the data-dependent count split at the fixed width b = 102 (normwise
    bound 4 * 2^-102 < 1e-30) and report that worst-case count; otherwise (and for complex
    products) b is at least 103."""
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE if estimate else mwblas.ESTIMATE_OFF)
    try:
        nmod, bits = mwblas.ozaki_config(k, cplx)
    finally:
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)
    mm, _ = mwblas.ozaki_moduli(nmod)
    M = math.prod(int(v) for v in mm)
    k_eff = 2 * k if cplx else k
    # |C'| <= k_eff * 2^(2b) must lie strictly inside the symmetric range (-M/2, M/2)
    assert M > 2 * k_eff * 2 ** (2 * bits)
    if estimate and not cplx:
        assert bits == 102
    else:
        assert bits >= 103          # normwise error bound 4 * 2^-b <= 4e-31
    assert 4 * 2.0 ** -bits < 1e-30
    # and one modulus fewer would not do
    if nmod > 1:
        M1 = M // int(mm[-1])
        assert not (M1 > 2 * k_eff * 2 ** (2 * min(bits, 103)))


def test_depth_limit_is_rejected():
    """The residue GEMMs reduce their sums modulo m every 32768 of depth, so the INT32
    accumulation sets no depth limit; the plans stop at k_eff = 2^24 (k real, 2k complex)."""
    mwblas.ozaki_config(1 << 24)
    mwblas.ozaki_config(1 << 23, cplx=True)
    with pytest.raises(mwblas.MwgemmError):
        mwblas.ozaki_config((1 << 24) + 1)
    with pytest.raises(mwblas.MwgemmError):
        mwblas.ozaki_config((1 << 23) + 1, cplx=True)


def test_explicit_exact_product_beyond_the_plans_fails_loudly():
    """An explicit exact-product request deeper than the plans cover is an error, not a
    silent plain-path run (AUTO takes the plain path there).  The depth is checked before
    any pointer is used or any device call is made."""
    k = (1 << 24) + 64
    for cplx in (False, True):
        w = 4 if cplx else 2
        with pytest.raises(mwblas.MwgemmError, match="depth"):
            mwblas.gemm_device(cplx, "T", "N", 4, 4, k, [1.0] + [0.0] * (w - 1), None, k, None,
                               k, [0.0] * w, None, 4, path=mwblas.PATH_OZAKI)


def test_fixed_point_crt_reconstruction_identity():
    """C'/M == frac(sum_t u_t y_t/m_t), u_t = C' mod m_t (what the residue GEMMs store),
    evaluated as sum_t u_t * W_t mod 2^128 with the weight-folded limbs W_t =
    round(2^128 y_t/m_t) mod 2^128: the error is below M * 2^-115 for every |C'| up to
    M/2 * (1 - 2^-100) (the plan keeps |C'| below M/2 by a factor 1 - 7e-7 or more).  The
    unfolded form (u_t = (C' mod m_t) y_t mod m_t with round(2^128/m_t)) is checked too:
    both give C' to the same bound, so dropping y_t from one side only would fail."""
    nmod = 29
    mm, yy = mwblas.ozaki_moduli(nmod)
    mm = [int(v) for v in mm]
    yy = [int(v) for v in yy]
    M = math.prod(mm)
    W = [(2 ** 128 + m // 2) // m for m in mm]
    Wy = [((2 ** 129 * y + m) // (2 * m)) % 2 ** 128 for m, y in zip(mm, yy)]
    rng = random.Random(5)
    lim = M // 2 - M // 2 ** 101

    def recon(u, weights):
        F = sum(ut * w for ut, w in zip(u, weights)) % 2 ** 128
        if F >= 2 ** 127:
            F -= 2 ** 128
        return F * M / 2 ** 128       # exact rational

    for _ in range(2000):
        X = rng.randrange(-lim, lim + 1)
        if rng.random() < 0.1:
            X = rng.choice([0, 1, -1, lim, -lim])
        plain = [X % m for m in mm]
        assert abs(recon(plain, Wy) - X) <= M / 2 ** 115
        assert abs(recon([(u * y) % m for u, y, m in zip(plain, yy, mm)], W) - X) <= M / 2 ** 115
        # any representative of the class will do: the symmetric residues the residue GEMMs
        # store, and the unreduced combinations of the complex reconstruction (|u| <= 384)
        sym = [((X + m // 2) % m) - m // 2 for m in mm]
        assert abs(recon(sym, Wy) - X) <= M / 2 ** 115
        shifted = [u + rng.choice((-1, 0, 1)) * m for u, m in zip(sym, mm)]
        assert abs(recon(shifted, Wy) - X) <= M / 2 ** 114
    # dropping the weight on one side is caught
    X = lim // 3
    assert abs(recon([X % m for m in mm], W) - X) > M / 2 ** 20


def _epilogue_residue(v, m):
    """The residue GEMMs' epilogue (crt_residue): t = fma(v, fl(1/m), 1.5 * 2^52) in binary64
    (the exact product plus the constant, rounded once), q = the low 32 bits of t as a signed
    integer, u = v - q * m with q inside the signed 24-bit range of the multiply.  Also
    returns the byte the stores take (crt_residue_byte): the low 8 bits of the 32-bit
    v + (low 24 bits of the word q) * (2^24 - m), one unsigned 24-bit multiply-add."""
    import struct
    from fractions import Fraction
    t = float(Fraction(v) * Fraction(1.0 / m) + Fraction(3, 2) * 2 ** 52)
    lo = struct.unpack("<q", struct.pack("<d", t))[0] & 0xFFFFFFFF
    q = lo - 2 ** 32 if lo >= 2 ** 31 else lo
    assert -2 ** 23 <= q < 2 ** 23, (v, m, q)
    byte = ((v & 0xFFFFFFFF) + (lo & 0xFFFFFF) * (2 ** 24 - m)) & 0xFF
    return v - q * m, byte


def test_epilogue_residue_is_a_symmetric_representative():
    """The residue GEMMs reduce every INT32 sum v (|v| <= 2^29 + 2^10: sums of up to 32768
    products of residues, plus a folded remainder below 4m) to v - m * round(v/m) with one
    FMA against 1.5 * 2^52 and one 24-bit multiply.  For every modulus: random sums, the
    range ends, and the integers next to every half-integer multiple of m near 0, the ends
    and random points; the result must be congruent to v modulo m and lie in the symmetric
    range (+-128 for m = 256), and the stored byte from the multiply-add form must be its
    low byte."""
    m_all, _ = mwblas.ozaki_moduli(40)
    rng = random.Random(17)
    lim = 2 ** 29 + 2 ** 10
    for m in (int(v) for v in m_all):
        top = lim // m
        h = m // 2
        cases = [rng.randrange(-lim, lim + 1) for _ in range(1500)]
        for q in (0, 1, top - 1, top) + tuple(rng.randrange(0, top) for _ in range(40)):
            cases += [q * m + d for d in (-h - 1, -h, -h + 1, -1, 0, 1, h - 1, h, h + 1)]
        cases += [-c for c in cases]
        bound = 128 if m == 256 else (m - 1) // 2
        for v in cases:
            if abs(v) > lim:
                continue
            u, byte = _epilogue_residue(v, m)
            assert (u - v) % m == 0 and -bound <= u <= bound, (m, v, u)
            assert byte == u & 0xFF, (m, v, u, byte)


def _sym_residue(s, m):
    """The split kernels' residue step in binary64: s is an exact integer (|s| < 2^45), the
    FMA after the quotient is an exact integer operation, only s * fl(1/m) rounds, and the
    quotient is rounded to the nearest integer (ties to even, as rint)."""
    q = round(float(s) * (1.0 / m))
    return s - q * m


def test_symmetric_residue_is_exact_for_both_signs():
    """r = s - m*rint(s*fl(1/m)) needs no offset and no fix-up: s*fl(1/m) is within
    2^-7/m of s/m, and for odd m every s/m is at least 1/(2m) from a half-integer, so r is
    the unique residue in [-(m-1)/2, (m-1)/2]; for m = 256 the product is exact and a tie
    gives +-128, whose byte is 0x80 either way.  Checked over the whole chunk range
    |s| < 2^45 (c0 < 2^35, |c1| < 2^36, |c2| <= 2^34, chunk weights below 256) for every
    modulus, with the integers next to every half-integer multiple of m, against the
    canonical residue's byte (the int8 the residue GEMM reads)."""
    m_all, _ = mwblas.ozaki_moduli(40)
    rng = random.Random(11)
    lim = 2 ** 45
    for m in (int(v) for v in m_all):
        top = (lim - 1) // m
        h = m // 2
        cases = [rng.randrange(-lim + 1, lim) for _ in range(2000)]
        cases += [q * m + d for q in (0, 1, 2, top - 1, top) for d in (-1, 0, 1, h, h + 1)]
        cases += [rng.randrange(0, top) * m + d for d in (-1, 0, 1, h, h + 1, m - 1)
                  for _ in range(100)]
        cases += [-c for c in cases]
        bound = 128 if m == 256 else (m - 1) // 2
        for s in cases:
            if abs(s) >= lim:
                continue
            r = _sym_residue(s, m)
            assert (r - s) % m == 0 and -bound <= r <= bound, (m, s, r)
            canon = s % m
            canon = canon - m if canon > (m - 1) // 2 else canon
            assert (r & 0xff) == (canon & 0xff), (m, s, r, canon)
            assert ((-r) & 0xff) == ((-canon) & 0xff), (m, s, r, canon)


def _chunks(hi, lo, sigma):
    """The split's integer chunking, step by step in binary64 (every step is exact)."""
    import numpy as np
    xh_s, xl_s = math.ldexp(hi, sigma), math.ldexp(lo, sigma)
    xh = float(np.rint(xh_s))
    xl = float(np.rint((xh_s - xh) + xl_s))
    ax = abs(xh)
    a2 = math.floor(ax * 2.0 ** -70)
    rem = ax - a2 * 2.0 ** 70
    a1 = math.floor(rem * 2.0 ** -35)
    a0 = rem - a1 * 2.0 ** 35
    sg = -1.0 if xh < 0 else 1.0
    t0 = sg * a0 + xl
    carry = math.floor(t0 * 2.0 ** -35)
    return t0 - carry * 2.0 ** 35, sg * a1 + carry, sg * a2, xh_s, xl_s


def test_split_chunks_are_exact_and_bounded():
    from fractions import Fraction
    import numpy as np
    import ddgen
    rng = np.random.default_rng(10)
    b = 103
    for kind in ("uniform", "wide"):
        vals = ddgen.real_values(rng, (4000,), kind=kind)
        mx = float(np.max(np.abs(vals[:, 0]) + np.abs(vals[:, 1])))
        sigma = b - math.frexp(mx)[1]
        for hi, lo in vals:
            c0, c1, c2, xh_s, xl_s = _chunks(float(hi), float(lo), sigma)
            assert 0 <= c0 < 2 ** 35 and abs(c1) < 2 ** 36 and abs(c2) <= 2 ** 34
            X = int(c0) + int(c1) * 2 ** 35 + int(c2) * 2 ** 70
            assert abs(X) <= 2 ** b
            assert abs(Fraction(X) - Fraction(xh_s) - Fraction(xl_s)) <= Fraction(5, 8)
            # the chunk combination stays below 2^45 for every modulus weight < 256
            assert abs(c0) + abs(c1) * 255 + abs(c2) * 255 < 2 ** 45


def test_residue_product_layout_decoder():
    """mwblas.residue_products decodes the MFMA block order of the residue products: the word
    of rows i..i+3 (i % 4 == 0) of column j sits at block (j // 32) * (mp // 32) + i // 32
    (1 KB each), word (i % 32) // 8 * 64 + (j % 32) + 32 * ((i // 4) % 2), byte i % 4.
    Each entry is placed by that rule and must come back at (j, i)."""
    import numpy as np
    import mwblas
    mp, np_ = 96, 64
    flat = np.zeros(mp * np_, dtype=np.int64)
    for j in range(np_):
        for i in range(mp):
            word = (i % 32) // 8 * 64 + (j % 32) + 32 * ((i // 4) % 2)
            off = ((j // 32) * (mp // 32) + i // 32) * 1024 + word * 4 + i % 4
            flat[off] = j * 1000 + i
    got = mwblas.residue_products(flat, mp, np_)
    expect = np.array([[j * 1000 + i for i in range(mp)] for j in range(np_)])
    assert np.array_equal(got, expect)
