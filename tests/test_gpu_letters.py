"""The drop-in entry points read trans and uplo letters outside N/T/C and U/L as the reference
does (its packers test for the letters they know and treat the rest alike; the readings are
pinned on the CPU restatement in tests/test_oracle_letters.py), bit for bit against the
restatement called with the same letters.

Real: any non-'N' letter transposes.  Complex op(A): anything but N/T conjugate-transposes;
op(B): the AVX2 kernels transpose without conjugation, the scalar kernels conjugate-transpose.
Complex gemmtr: the diagonal triangles read such a letter as 'T', the rectangle as the GEMM
does.  gemmtr uplo: anything but 'U' is the lower triangle.  Complex gemv: anything but N/C
transposes."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle
from test_gpu_gemv import _dropin_bands

pytestmark = pytest.mark.gpu

ORDERS = [(mwblas.ORDER_SIMD, oracle.ORDER_SIMD), (mwblas.ORDER_SCALAR, oracle.ORDER_SCALAR)]


def _with_order(order, fn):
    prev_g = mwblas.set_plain_order(order[0])
    prev_p = oracle.set_order(order[1])
    try:
        return fn()
    finally:
        mwblas.set_plain_order(prev_g)
        oracle.set_order(prev_p)


@pytest.mark.parametrize("order", ORDERS, ids=["simd", "scalar"])
def test_gemm_letters_bitwise(gpu, order):
    rng = np.random.default_rng(901 + order[0])
    m, n, k = 77, 45, 300

    def run():
        for cplx, ta, tb in ((False, "x", "q"), (False, "n", "c"), (False, "*", "N"),
                             (True, "X", "X"), (True, "n", "z"), (True, "c", "t"),
                             (True, "q", "c")):
            gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
            scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
            ar, ac = (m, k) if ta.upper() == "N" else (k, m)
            br, bc = (k, n) if tb.upper() == "N" else (n, k)
            A, B, C0 = gen(rng, ar, ac), gen(rng, br, bc), gen(rng, m, n)
            al, be = scal(rng, scale=2), scal(rng, scale=2)
            g, p = C0.copy(), C0.copy()
            (mwblas.wgemm if cplx else mwblas.mgemm)(ta, tb, m, n, k, al, A, ar, B, br, be, g, m)
            (oracle.port_wgemm if cplx else oracle.port_mgemm)(ta, tb, m, n, k, al, A, ar, B,
                                                                br, be, p, m)
            assert np.array_equal(g, p), (cplx, ta, tb)
            g2 = C0.copy()
            (mwblas.wgemm_serial if cplx else mwblas.mgemm_serial)(ta, tb, m, n, k, al, A, ar,
                                                                    B, br, be, g2, m)
            assert np.array_equal(g2, p), ("serial", cplx, ta, tb)

    _with_order(order, run)


@pytest.mark.parametrize("order", ORDERS, ids=["simd", "scalar"])
def test_gemmtr_letters_bitwise(gpu, order):
    rng = np.random.default_rng(911 + order[0])
    n, k = 133, 90

    def run():
        for cplx, uplo, ta, tb in ((False, "x", "q", "z"), (False, "u", "n", "c"),
                                   (True, "w", "X", "x"), (True, "u", "X", "n"),
                                   (True, "U", "n", "y"), (True, "l", "c", "q")):
            gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
            scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
            ar, ac = (n, k) if ta.upper() == "N" else (k, n)
            br, bc = (k, n) if tb.upper() == "N" else (n, k)
            A, B, C0 = gen(rng, ar, ac), gen(rng, br, bc), gen(rng, n, n)
            al, be = scal(rng, scale=2), scal(rng, scale=2)
            g, p = C0.copy(), C0.copy()
            (mwblas.wgemmtr if cplx else mwblas.mgemmtr)(uplo, ta, tb, n, k, al, A, ar, B, br,
                                                         be, g, n)
            (oracle.port_wgemmtr if cplx else oracle.port_mgemmtr)(uplo, ta, tb, n, k, al, A,
                                                                   ar, B, br, be, p, n)
            assert np.array_equal(g, p), (cplx, uplo, ta, tb)

    _with_order(order, run)


def test_gemv_letters_bitwise(gpu):
    rng = np.random.default_rng(921)
    m, n = 203, 150
    for cplx, trans in ((False, "x"), (False, "c"), (True, "x"), (True, "c"), (True, "t")):
        gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
        scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
        notrans = trans.upper() == "N"
        A = gen(rng, m, n)
        x = gen(rng, n if notrans else m, 1)[0]
        y0 = gen(rng, m if notrans else n, 1)[0]
        al, be = scal(rng, scale=2), scal(rng, scale=2)
        g, p = y0.copy(), y0.copy()
        (mwblas.wgemv if cplx else mwblas.mgemv)(trans, m, n, al, A, m, x, 1, be, g, 1)
        if cplx:
            oracle.port_wgemv(trans, m, n, al, A, m, x, 1, be, p, 1)
        else:
            oracle.port_mgemv(trans, m, n, al, A, m, x, 1, be, p, 1,
                              nbands=_dropin_bands(m) if notrans else 1)
        assert np.array_equal(g, p), (cplx, trans)


def test_exact_product_path_reads_letters_alike(gpu):
    """The exact-product path resolves the letters the same way: a complex ('X','X') call
    is within the bound of the binary128 product taken with the letters the plain path
    reads, ('C','T') in the AVX2 order."""
    rng = np.random.default_rng(931)
    m, n, k = 96, 80, 200
    A, B, C0 = ddgen.cplx_matrix(rng, k, m), ddgen.cplx_matrix(rng, n, k), ddgen.cplx_matrix(rng, m, n)
    al, be = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)
    prev_o = mwblas.set_plain_order(mwblas.ORDER_SIMD)
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    try:
        g = C0.copy()
        mwblas.wgemm("X", "x", m, n, k, al, A, k, B, n, be, g, m)
    finally:
        mwblas.set_path(prev)
        mwblas.set_plain_order(prev_o)
    r = oracle.q128_gemm_err(True, "C", "T", m, n, k, al, A, k, B, n, be, C0, m, g, m)
    assert r["err"] <= 1e-30, r
    wrong = oracle.q128_gemm_err(True, "C", "C", m, n, k, al, A, k, B, n, be, C0, m, g, m)
    assert wrong["err"] > 1e-20
