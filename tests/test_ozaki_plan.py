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
                                    (4096, True), (32768, True)])
def test_plan_covers_the_product_range(k, cplx):
    nmod, bits = mwblas.ozaki_config(k, cplx)
    mm, _ = mwblas.ozaki_moduli(nmod)
    M = math.prod(int(v) for v in mm)
    k_eff = 2 * k if cplx else k
    # |C'| <= k_eff * 2^(2b) must lie strictly inside the symmetric range (-M/2, M/2)
    assert M > 2 * k_eff * 2 ** (2 * bits)
    assert bits >= 103          # normwise error bound 4 * 2^-b <= 4e-31
    # and one modulus fewer would not do
    if nmod > 1:
        M1 = M // int(mm[-1])
        assert not (M1 > 2 * k_eff * 2 ** (2 * 103))


def test_depth_limit_is_rejected():
    with pytest.raises(mwblas.MwgemmError):
        mwblas.ozaki_config(65537)
    with pytest.raises(mwblas.MwgemmError):
        mwblas.ozaki_config(32769, cplx=True)


def test_fixed_point_crt_reconstruction_identity():
    """C'/M == frac(sum_t u_t/m_t), u_t = (C' mod m_t) * y_t mod m_t, evaluated as
    sum_t u_t * round(2^128/m_t) mod 2^128: the error is below M * 2^-115 for every |C'|
    up to M/2 * (1 - 2^-100) (the plan keeps |C'| below M/2 by a factor 1 - 7e-7 or more)."""
    nmod = 29
    mm, yy = mwblas.ozaki_moduli(nmod)
    mm = [int(v) for v in mm]
    yy = [int(v) for v in yy]
    M = math.prod(mm)
    W = [(2 ** 128 + m // 2) // m for m in mm]
    rng = random.Random(5)
    lim = M // 2 - M // 2 ** 101
    for _ in range(2000):
        X = rng.randrange(-lim, lim + 1)
        if rng.random() < 0.1:
            X = rng.choice([0, 1, -1, lim, -lim])
        u = [((X % m) * y) % m for m, y in zip(mm, yy)]
        F = sum(ut * w for ut, w in zip(u, W)) % 2 ** 128
        if F >= 2 ** 127:
            F -= 2 ** 128
        est = F * M / 2 ** 128       # exact rational
        assert abs(est - X) <= M / 2 ** 115
