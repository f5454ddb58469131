"""Pin how the CPU restatement reads trans and uplo letters outside N/T/C and U/L.

The reference never rejects a letter.  Each of its packers tests for the letters it knows
and treats every other letter like one of them, so a letter such as 'X' computes with a
definite operation:
* real GEMM, gemmtr and gemv: blas_trans_real upper-cases and folds 'C' to 'T'
  (common/blas_char.h:33-36), then the packers test == 'N' (mgemm_serial.cpp:52, 71;
  mgemv.cpp:248-250), so every other letter is a transpose;
* complex op(A): wgemm_pack_A handles 'N' and 'T' and conjugate-transposes the rest
  (wgemm_serial.cpp:56-73);
* complex op(B): the AVX2 packer transposes every non-'N' letter and conjugates only for
  'C' (wgemm_serial.cpp:145, 174-185); the scalar packer handles 'N' and 'T' and
  conjugate-transposes the rest (:81-99);
* complex gemmtr diagonal triangles: transposed unless 'N', conjugated only for 'C'
  (wgemmtr_serial.cpp:54-57), while the rectangle goes through the GEMM packers above;
* gemmtr uplo: 'U' is the upper triangle, anything else the lower (mgemmtr_serial.cpp:178);
* complex gemv: 'N', 'C', anything else transposed (wgemv.cpp:308-319).
Each case is pinned by bitwise equality with the same call made with the standard letter the
reference's code reads it as, so a restatement that read 'X' as the wrong operation (or
ignored case) fails here.
"""
import numpy as np
import pytest

import ddgen
import oracle


def _real_case(rng, m, n, k):
    A = ddgen.real_matrix(rng, k, m)          # k x m: op(A) transposed
    B = ddgen.real_matrix(rng, n, k)          # n x k: op(B) transposed
    C0 = ddgen.real_matrix(rng, m, n)
    return A, B, C0, ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)


def _cplx_case(rng, m, n, k):
    A = ddgen.cplx_matrix(rng, k, m)
    B = ddgen.cplx_matrix(rng, n, k)
    C0 = ddgen.cplx_matrix(rng, m, n)
    return A, B, C0, ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)


def _mgemm(ta, tb, m, n, k, case):
    A, B, C0, al, be = case
    C = C0.copy()
    oracle.port_mgemm(ta, tb, m, n, k, al, A, k, B, n, be, C, m)
    return C


def _wgemm(ta, tb, m, n, k, case):
    A, B, C0, al, be = case
    C = C0.copy()
    oracle.port_wgemm(ta, tb, m, n, k, al, A, k, B, n, be, C, m)
    return C


@pytest.mark.parametrize("order", [oracle.ORDER_SIMD, oracle.ORDER_SCALAR], ids=["simd", "scalar"])
def test_real_gemm_reads_any_non_n_letter_as_transpose(order):
    rng = np.random.default_rng(501)
    m, n, k = 37, 29, 300
    case = _real_case(rng, m, n, k)
    prev = oracle.set_order(order)
    try:
        want = _mgemm("T", "T", m, n, k, case)
        for ta, tb in (("X", "q"), ("c", "C"), ("t", "z"), ("*", "T")):
            assert np.array_equal(_mgemm(ta, tb, m, n, k, case), want), (ta, tb)
    finally:
        oracle.set_order(prev)


def test_complex_gemm_letters_simd_order():
    rng = np.random.default_rng(502)
    m, n, k = 33, 21, 270
    case = _cplx_case(rng, m, n, k)
    prev = oracle.set_order(oracle.ORDER_SIMD)
    try:
        # op(A): an unknown letter conjugates; op(B): the AVX2 packer only transposes
        want = _wgemm("C", "T", m, n, k, case)
        for ta, tb in (("X", "X"), ("c", "q"), ("z", "t")):
            assert np.array_equal(_wgemm(ta, tb, m, n, k, case), want), (ta, tb)
        assert not np.array_equal(want, _wgemm("T", "T", m, n, k, case))
        assert not np.array_equal(want, _wgemm("C", "C", m, n, k, case))
    finally:
        oracle.set_order(prev)


def test_complex_gemm_letters_scalar_order():
    rng = np.random.default_rng(503)
    m, n, k = 19, 23, 150
    case = _cplx_case(rng, m, n, k)
    prev = oracle.set_order(oracle.ORDER_SCALAR)
    try:
        # both scalar packers conjugate-transpose an unknown letter
        want = _wgemm("C", "C", m, n, k, case)
        for ta, tb in (("X", "X"), ("c", "y"), ("q", "c")):
            assert np.array_equal(_wgemm(ta, tb, m, n, k, case), want), (ta, tb)
        assert not np.array_equal(want, _wgemm("C", "T", m, n, k, case))
    finally:
        oracle.set_order(prev)


def test_gemmtr_uplo_other_than_u_is_lower():
    rng = np.random.default_rng(504)
    n, k = 45, 70
    A = ddgen.real_matrix(rng, n, k)
    B = ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, n, n)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)

    def run(uplo, ta="N", tb="N"):
        C = C0.copy()
        oracle.port_mgemmtr(uplo, ta, tb, n, k, al, A, n, B, k, be, C, n)
        return C

    lower, upper = run("L"), run("U")
    assert not np.array_equal(lower, upper)
    for u in ("l", "X", "0"):
        assert np.array_equal(run(u), lower), u
    assert np.array_equal(run("u"), upper)


def _diag_mask(n, nb, upper):
    """Entries of the nb x nb diagonal triangles (the reference's scalar part)."""
    mask = np.zeros((n, n), dtype=bool)          # [col, row]
    for jc in range(0, n, nb):
        jb = min(nb, n - jc)
        for j in range(jc, jc + jb):
            rows = range(jc, j + 1) if upper else range(j, jc + jb)
            for i in rows:
                mask[j, i] = True
    return mask


@pytest.mark.parametrize("order", [oracle.ORDER_SIMD, oracle.ORDER_SCALAR], ids=["simd", "scalar"])
@pytest.mark.parametrize("upper", [True, False], ids=["U", "L"])
def test_complex_gemmtr_unknown_letters_split_by_part(order, upper):
    """With ('X','X') the diagonal triangles compute as ('T','T') while the rectangle
    computes as the GEMM reads the letters: ('C','T') in the AVX2 order, ('C','C') in the
    scalar order."""
    rng = np.random.default_rng(505 + 2 * upper + order)
    n, k, nb = 41, 60, 8
    A = ddgen.cplx_matrix(rng, k, n)
    B = ddgen.cplx_matrix(rng, n, k)
    C0 = ddgen.cplx_matrix(rng, n, n)
    al, be = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)
    uplo = "U" if upper else "L"

    def run(ta, tb):
        C = C0.copy()
        oracle.port_wgemmtr(uplo, ta, tb, n, k, al, A, k, B, n, be, C, n, nb=nb)
        return C

    prev = oracle.set_order(order)
    try:
        got = run("X", "x")
        rect = run("C", "T" if order == oracle.ORDER_SIMD else "C")
        diag = run("T", "T")
    finally:
        oracle.set_order(prev)
    mask = _diag_mask(n, nb, upper)
    assert np.array_equal(got[mask], diag[mask])
    assert np.array_equal(got[~mask], rect[~mask])
    assert not np.array_equal(rect[mask], diag[mask])


def test_gemv_letters():
    rng = np.random.default_rng(507)
    m, n = 70, 50
    A = ddgen.real_matrix(rng, m, n)
    x = ddgen.real_matrix(rng, m, 1)[0]
    y0 = ddgen.real_matrix(rng, n, 1)[0]
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)

    def rrun(t):
        y = y0.copy()
        oracle.port_mgemv(t, m, n, al, A, m, x, 1, be, y, 1)
        return y

    for t in ("x", "c", "Q"):
        assert np.array_equal(rrun(t), rrun("T")), t
    Z = ddgen.cplx_matrix(rng, m, n)
    zx = ddgen.cplx_matrix(rng, m, 1)[0]
    zy0 = ddgen.cplx_matrix(rng, n, 1)[0]
    zal, zbe = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)

    def zrun(t):
        y = zy0.copy()
        oracle.port_wgemv(t, m, n, zal, Z, m, zx, 1, zbe, y, 1)
        return y

    # 'N' and 'C' are the only letters wgemv tests for; everything else transposes
    for t in ("x", "t", "*"):
        assert np.array_equal(zrun(t), zrun("T")), t
    assert np.array_equal(zrun("c"), zrun("C"))
    assert not np.array_equal(zrun("C"), zrun("T"))
