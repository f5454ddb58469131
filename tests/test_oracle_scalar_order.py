"""Pin the CPU restatement of the reference's scalar kernels, the operation order it runs on
hosts without AVX2/FMA (mgemm_serial.cpp:87-99, wgemm_serial.cpp:102-115, mgemv.cpp:147-153
and :213-223, wgemv.cpp:170-188 and :284-295; chosen by mf_have_avx2_fma, mf_cpu.cpp:10-23).

Per element of C: C := beta-prepass(C), then for p ascending C := C (+) ((alpha (x) b_pj)
(x) a_ip), faithful operations, no 128-deep chunks.  What pins it:
* no chunk structure: splitting K at ANY depth into two calls (beta = 1 on the second) is
  bit-identical to one call (the SIMD order fails this at depths that are not multiples of
  128, which the test also shows);
* k = 1 reduces the whole sequence to one step, C := prepass(C0) (+) ((alpha (x) b) (x) a),
  checked through the separately pinned faithful primitives;
* the reference fuzz tolerance against binary128, exact integer products, transposition and
  conjugation invariance;
* mgemv: op 'N' does not depend on the row-band count (no 4-blocks, no bands) and equals
  the GEMM with x as the single column; op 'T' with alpha = 1, beta = 0 equals the GEMM
  ('T','N', n x 1 x m) on the same data: two code paths, one sequence.
"""
import ctypes

import numpy as np
import pytest

import ddgen
import oracle

TOL_EPS = 2.0 ** -104


@pytest.fixture
def scalar_order():
    prev = oracle.set_order(oracle.ORDER_SCALAR)
    yield
    oracle.set_order(prev)


def at(X, i, j):
    """Interior pointer to element (i, j) of a column-major DD array (cols, ld, limbs)."""
    return X.reshape(-1, X.shape[2])[j * X.shape[1] + i:]


def _prim(name, a, b):
    out = np.zeros(2)
    aa, bb = np.ascontiguousarray(a, float), np.ascontiguousarray(b, float)
    getattr(oracle.lib(), name)(aa.ctypes.data_as(ctypes.c_void_p),
                                bb.ctypes.data_as(ctypes.c_void_p),
                                out.ctypes.data_as(ctypes.c_void_p))
    return out


def _split_equal(cplx, ta, tb, m, n, k, k1, rng):
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A, B = gen(rng, ar, ac, ar + 2), gen(rng, br, bc, br + 1)
    C0 = gen(rng, m, n, m + 3)
    scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
    al, be = scal(rng, scale=2), scal(rng, scale=2)
    one = np.array([1.0, 0, 0, 0]) if cplx else np.array([1.0, 0])
    port = oracle.port_wgemm if cplx else oracle.port_mgemm
    C1 = C0.copy()
    port(ta, tb, m, n, k, al, A, ar + 2, B, br + 1, be, C1, m + 3)
    C2 = C0.copy()
    port(ta, tb, m, n, k1, al, A, ar + 2, B, br + 1, be, C2, m + 3)
    a2 = at(A, 0, k1) if ta == "N" else at(A, k1, 0)
    b2 = at(B, k1, 0) if tb == "N" else at(B, 0, k1)
    port(ta, tb, m, n, k - k1, al, a2, ar + 2, b2, br + 1, one, C2, m + 3)
    return np.array_equal(C1, C2, equal_nan=True)


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_k_split_at_any_depth(scalar_order, cplx):
    rng = np.random.default_rng(900 + cplx)
    ops = [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")] + ([("C", "N"), ("N", "C")] if cplx else [])
    for ta, tb in ops:
        for k1 in (1, 37, 64, 100, 149):
            assert _split_equal(cplx, ta, tb, 23, 17, 150, k1, rng), (ta, tb, k1)


def test_simd_order_is_not_split_invariant():
    """The same property fails in the SIMD order at depth 37 (chunks of 128): the two
    orders are distinct sequences and the test above tells them apart."""
    rng = np.random.default_rng(910)
    assert oracle.lib().port_get_order() == oracle.ORDER_SIMD
    assert not _split_equal(False, "N", "N", 23, 17, 150, 37, rng)


def test_k1_is_one_faithful_step(scalar_order):
    rng = np.random.default_rng(920)
    for trial in range(20):
        m, n = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        A, B, C0 = (ddgen.real_matrix(rng, m, 1), ddgen.real_matrix(rng, 1, n),
                    ddgen.real_matrix(rng, m, n))
        al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
        if trial % 5 == 1:
            be = np.array([1.0, 0.0])
        C = C0.copy()
        oracle.port_mgemm("N", "N", m, n, 1, al, A, m, B, 1, be, C, m)
        for j in range(n):
            t = _prim("port_dd_mul_exact", al, B[j, 0])
            for i in range(m):
                c = C0[j, i] if trial % 5 == 1 else _prim("port_dd_mul_exact", C0[j, i], be)
                expect = _prim("port_dd_add_exact", c, _prim("port_dd_mul_exact", t, A[0, i]))
                assert np.array_equal(C[j, i], expect), (trial, i, j)


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_fuzz_against_binary128(scalar_order, cplx):
    rng = np.random.default_rng(930 + cplx)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    letters = "NTC" if cplx else "NT"
    for _ in range(30):
        m, n, k = (ddgen.rand_int_log(rng, 1, 200) for _ in range(3))
        ta, tb = letters[int(rng.integers(0, len(letters)))], letters[int(rng.integers(0, len(letters)))]
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        A, B, C0 = gen(rng, ar, ac, ar + 1), gen(rng, br, bc, br + 2), gen(rng, m, n, m + 1)
        al, be = ddgen.fuzz_alpha_beta(rng, cplx)
        C = C0.copy()
        (oracle.port_wgemm if cplx else oracle.port_mgemm)(ta, tb, m, n, k, al, A, ar + 1, B,
                                                          br + 2, be, C, m + 1)
        r = oracle.q128_gemm_err(cplx, ta, tb, m, n, k, al, A, ar + 1, B, br + 2, be, C0, m + 1,
                                 C, m + 1)
        assert r["err"] <= max(10, 4 * k) * TOL_EPS, (ta, tb, m, n, k, r)


def test_integer_operands_and_transposition(scalar_order):
    rng = np.random.default_rng(940)
    m, n, k = 31, 19, 77
    A = ddgen.real_matrix(rng, m, k, kind="integer", bound=512)
    B = ddgen.real_matrix(rng, k, n, kind="integer", bound=512)
    C0 = ddgen.real_matrix(rng, m, n, kind="integer", bound=1000)
    al, be = np.array([3.0, 0.0]), np.array([-2.0, 0.0])
    C = C0.copy()
    oracle.port_mgemm("N", "N", m, n, k, al, A, m, B, k, be, C, m)
    ia, ib = A[..., 0].T.astype(np.int64), B[..., 0].T.astype(np.int64)
    exact = 3 * (ia @ ib) - 2 * C0[..., 0].T.astype(np.int64)
    assert np.array_equal(C[..., 0].T, exact.astype(np.float64)) and np.all(C[..., 1] == 0)
    # NN == TN(A^T) == NT(B^T) == TT bitwise, on random data
    A, B = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n)
    At = np.ascontiguousarray(np.transpose(A, (1, 0, 2)))
    Bt = np.ascontiguousarray(np.transpose(B, (1, 0, 2)))
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    outs = []
    for ta, tb, AA, lda, BB, ldb in [("N", "N", A, m, B, k), ("T", "N", At, k, B, k),
                                     ("N", "T", A, m, Bt, n), ("T", "T", At, k, Bt, n)]:
        C = C0.copy()
        oracle.port_mgemm(ta, tb, m, n, k, al, AA, lda, BB, ldb, be, C, m)
        outs.append(C)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_complex_conjugation_folding(scalar_order):
    rng = np.random.default_rng(950)
    m, n, k = 13, 11, 40
    A, B = ddgen.cplx_matrix(rng, k, m), ddgen.cplx_matrix(rng, k, n)
    C0 = ddgen.cplx_matrix(rng, m, n)
    al, be = ddgen.cplx_scalar(rng, scale=2), ddgen.cplx_scalar(rng, scale=2)
    C1 = C0.copy()
    oracle.port_wgemm("C", "N", m, n, k, al, A, k, B, k, be, C1, m)
    Act = np.ascontiguousarray(np.transpose(A, (1, 0, 2)) * np.array([1, 1, -1, -1]))
    C2 = C0.copy()
    oracle.port_wgemm("N", "N", m, n, k, al, Act, m, B, k, be, C2, m)
    assert np.array_equal(C1, C2)


def test_gemv_scalar_order_against_the_gemm(scalar_order):
    rng = np.random.default_rng(960)
    m, n = 75, 58
    A = ddgen.real_matrix(rng, m, n)
    x = ddgen.real_matrix(rng, n, 1)[0]
    y0 = ddgen.real_matrix(rng, m, 1)[0]
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    ys = []
    for nb in (1, 3, 16):
        y = y0.copy()
        oracle.port_mgemv("N", m, n, al, A, m, x, 1, be, y, 1, nb)
        ys.append(y)
    assert all(np.array_equal(y, ys[0]) for y in ys[1:])
    C = y0.copy()[None]
    oracle.port_mgemm("N", "N", m, 1, n, al, A, m, x[None], n, be, C, m)
    assert np.array_equal(C[0], ys[0])
    # op 'T', alpha = 1, beta = 0: one sequential dot per column == GEMM ('T','N')
    xm = ddgen.real_matrix(rng, m, 1)[0]
    y = np.zeros((n, 2))
    oracle.port_mgemv("T", m, n, np.array([1.0, 0]), A, m, xm, 1, np.zeros(2), y, 1)
    C = np.zeros((1, n, 2))
    oracle.port_mgemm("T", "N", n, 1, m, np.array([1.0, 0]), A, m, xm[None], m, np.zeros(2), C, n)
    assert np.array_equal(C[0], y)
