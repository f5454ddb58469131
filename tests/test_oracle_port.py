"""Pin the CPU restatement of the reference mgemm_/wgemm_ (oracle/ddgemm_port.c).

What pins what (the reference ships no golden vectors for this path and its multifloats
dependency is absent, so the pins come from mathematics and from the reference's structure):

* error-free transforms: TwoProd / TwoSum outputs sum exactly to the exact product / sum
  (checked in binary128, exact for these widths);
* DD product / sum accuracy: the published error bounds of the algorithms
  (DWTimesDW with FMA <= 5u^2 relative; AccurateDWPlusDW <= 3u^2 relative; SloppyDWPlusDW
  <= 3u^2 relative for same-sign operands, Joldes-Muller-Popescu, ACM TOMS 44(2) 2017),
  which a dropped cross term or a wrong operand fails;
* whole GEMM: the reference fuzz tolerance max(10,4K)*2^-104 normwise against the binary128
  oracle (fuzz_mgemm.fypp:51-54, fuzz_util.fypp:252-284) over fuzz-drawn cases;
* exactness: integer operands make every DD op exact, so C equals the integer product;
* structure: transposition is pure data movement (NN == TN(A^T) == NT(B^T) bitwise, CN ==
  NN(conj A^T)); C is updated once per 128-deep K chunk from k = 0
  (mgemm_serial.cpp:35-37, :126, :173-191), so a K = 128+r product equals a 128 product
  followed by an r product with beta = 1, bit for bit, while splitting at 64 does not;
* special cases of the driver (mgemm_parallel.cpp:66-78).
Bitwise parity with the third-party multifloats scalar operators themselves is unpinned.
"""
import ctypes

import numpy as np
import pytest

import ddgen
import oracle

TOL_EPS = 2.0 ** -104
U2 = 2.0 ** -106


def _dd_array(rng, n, kind="uniform", **kw):
    return np.ascontiguousarray(ddgen.real_values(rng, (n,), kind=kind, **kw))


def test_two_prod_two_sum_are_error_free():
    rng = np.random.default_rng(1)
    L = oracle.lib()
    p, e = ctypes.c_double(), ctypes.c_double()
    for _ in range(3000):
        x = float(np.ldexp(rng.uniform(-2, 2), int(rng.integers(-60, 60))))
        y = float(np.ldexp(rng.uniform(-2, 2), int(rng.integers(-60, 60))))
        L.port_eft_prod(x, y, ctypes.byref(p), ctypes.byref(e))
        assert L.q128_check_prod(x, y, p.value, e.value)
        L.port_eft_sum(x, y, ctypes.byref(p), ctypes.byref(e))
        assert L.q128_check_sum(x, y, p.value, e.value)


@pytest.mark.parametrize("name,bound", [("port_dd_mul_fast", 5 * U2), ("port_dd_mul_exact", 7 * U2)])
def test_dd_products_within_published_bound(name, bound):
    rng = np.random.default_rng(2)
    L = oracle.lib()
    f = getattr(L, name)
    a = _dd_array(rng, 4000, kind="wide")
    b = _dd_array(rng, 4000, kind="wide")
    r = np.zeros(2)
    worst = 0.0
    for i in range(a.shape[0]):
        x, y = np.ascontiguousarray(a[i]), np.ascontiguousarray(b[i])
        f(x.ctypes.data, y.ctypes.data, r.ctypes.data)
        worst = max(worst, L.q128_dd_mul_relerr(x.ctypes.data, y.ctypes.data, r.ctypes.data))
    assert worst <= bound, worst
    assert worst > 0.0   # the products are not exact: the check is live


def test_dd_sums_within_published_bound():
    rng = np.random.default_rng(3)
    L = oracle.lib()
    a = _dd_array(rng, 4000, kind="wide")
    b = _dd_array(rng, 4000, kind="wide")
    r = np.zeros(2)
    for i in range(a.shape[0]):
        x, y = np.ascontiguousarray(a[i]), np.ascontiguousarray(b[i])
        exact_mag = abs((x[0] + y[0]))
        L.port_dd_add_exact(x.ctypes.data, y.ctypes.data, r.ctypes.data)
        err = L.q128_dd_add_abserr(x.ctypes.data, y.ctypes.data, r.ctypes.data)
        assert err <= 3.5 * U2 * max(exact_mag, abs(r[0])) + 1e-300
        if np.sign(x[0]) == np.sign(y[0]):
            L.port_dd_add_fast(x.ctypes.data, y.ctypes.data, r.ctypes.data)
            err = L.q128_dd_add_abserr(x.ctypes.data, y.ctypes.data, r.ctypes.data)
            assert err <= 3.5 * U2 * exact_mag


def _fuzz_case(rng, cplx):
    m, n, k = (ddgen.rand_int_log(rng, 1, 96) for _ in range(3))
    letters = "NTC"
    ta, tb = letters[int(rng.integers(0, 3))], letters[int(rng.integers(0, 3))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), m + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    A, B, C0 = gen(rng, ar, ac, lda), gen(rng, br, bc, ldb), gen(rng, m, n, ldc)
    return dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, A=A, lda=lda, B=B,
                ldb=ldb, C0=C0, ldc=ldc)


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_port_fuzz_against_binary128(cplx):
    rng = np.random.default_rng(42 + cplx)
    run = oracle.port_wgemm if cplx else oracle.port_mgemm
    for _ in range(40):
        c = _fuzz_case(rng, cplx)
        C = c["C0"].copy()
        run(c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
            c["ldb"], c["beta"], C, c["ldc"])
        r = oracle.q128_gemm_err(cplx, c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"],
                                 c["A"], c["lda"], c["B"], c["ldb"], c["beta"], c["C0"],
                                 c["ldc"], C, c["ldc"])
        tol = max(10, 4 * c["k"]) * TOL_EPS
        assert r["err"] <= tol, (c["ta"], c["tb"], c["m"], c["n"], c["k"], r)
        # padding rows of C are never written (sentinel check, fuzz_util.fypp:234-250)
        assert np.array_equal(C[:, c["m"]:, :], c["C0"][:, c["m"]:, :])


def test_integer_operands_give_the_exact_integer_product():
    rng = np.random.default_rng(5)
    for (m, n, k) in [(17, 9, 300), (64, 64, 257), (5, 70, 1)]:
        A = ddgen.real_matrix(rng, m, k, kind="integer", bound=64)
        B = ddgen.real_matrix(rng, k, n, kind="integer", bound=64)
        C0 = ddgen.real_matrix(rng, m, n, kind="integer", bound=1000)
        C = C0.copy()
        alpha, beta = np.array([1.0, 0.0]), np.array([-3.0, 0.0])
        oracle.port_mgemm("N", "N", m, n, k, alpha, A, m, B, k, beta, C, m)
        ia = A[..., 0].T.astype(np.int64)      # (m, k)
        ib = B[..., 0].T.astype(np.int64)      # (k, n)
        exact = ia @ ib - 3 * C0[..., 0].T.astype(np.int64)
        assert np.array_equal(C[..., 0].T, exact.astype(np.float64))
        assert np.all(C[..., 1] == 0.0)


def _transpose(X, rows, cols):
    """(cols, ld, w) column-major rows x cols -> its transpose, tight."""
    return np.ascontiguousarray(np.transpose(X[:, :rows, :], (1, 0, 2)))


def test_transposition_is_bitwise_neutral():
    rng = np.random.default_rng(6)
    m, n, k = 37, 29, 300
    A = ddgen.real_matrix(rng, m, k)
    B = ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    At, Bt = _transpose(A, m, k), _transpose(B, k, n)
    outs = []
    for ta, tb, a, lda, b, ldb in [("N", "N", A, m, B, k), ("T", "N", At, k, B, k),
                                   ("N", "T", A, m, Bt, n), ("C", "C", At, k, Bt, n)]:
        C = C0.copy()
        oracle.port_mgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, C, m)
        outs.append(C)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_conjugate_transpose_is_bitwise_neutral():
    rng = np.random.default_rng(7)
    m, n, k = 21, 13, 140
    A = ddgen.cplx_matrix(rng, k, m)          # stored k x m, used as A^H
    B = ddgen.cplx_matrix(rng, k, n)
    C0 = ddgen.cplx_matrix(rng, m, n)
    al, be = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)
    Ah = _transpose(A, k, m)
    Ah[..., 2:4] *= -1.0
    C1, C2 = C0.copy(), C0.copy()
    oracle.port_wgemm("C", "N", m, n, k, al, A, k, B, k, be, C1, m)
    oracle.port_wgemm("N", "N", m, n, k, al, Ah, m, B, k, be, C2, m)
    assert np.array_equal(C1, C2)


def test_k_chunk_flush_period_is_128():
    rng = np.random.default_rng(8)
    m, n, k = 19, 11, 300
    A = ddgen.real_matrix(rng, m, k)
    B = ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    one = np.array([1.0, 0.0])
    full = C0.copy()
    oracle.port_mgemm("N", "N", m, n, k, al, A, m, B, k, be, full, m)

    def split(at):
        C = C0.copy()
        oracle.port_mgemm("N", "N", m, n, at, al, A, m, B, k, be, C, m)
        A2 = np.ascontiguousarray(A[at:])
        B2 = np.ascontiguousarray(B[:, at:, :])
        oracle.port_mgemm("N", "N", m, n, k - at, al, A2, m, B2, k - at, one, C, m)
        return C

    assert np.array_equal(split(128), full)
    assert np.array_equal(split(256), full)
    assert not np.array_equal(split(64), full)


@pytest.mark.parametrize("ta", ["N", "C"])
def test_complex_integer_operands_give_the_exact_integer_product(ta):
    """Integer real and imaginary parts: every complex DD product and sum is exact, so C is
    the exact Gaussian-integer result of alpha op(A) B + beta C0 (a dropped cross term, a
    wrong sign in the imaginary part or a missing conjugation fails)."""
    rng = np.random.default_rng(15)
    prev = oracle.set_order(oracle.ORDER_SIMD)
    try:
        for (m, n, k) in [(13, 9, 300), (32, 33, 257), (5, 7, 1)]:
            ar, ac = (m, k) if ta == "N" else (k, m)
            A = ddgen.cplx_matrix(rng, ar, ac, kind="integer", bound=64)
            B = ddgen.cplx_matrix(rng, k, n, kind="integer", bound=64)
            C0 = ddgen.cplx_matrix(rng, m, n, kind="integer", bound=1000)
            alpha, beta = np.array([2.0, 0.0, -1.0, 0.0]), np.array([-3.0, 0.0, 1.0, 0.0])
            C = C0.copy()
            oracle.port_wgemm(ta, "N", m, n, k, alpha, A, ar, B, k, beta, C, m)
            za = (A[..., 0] + 1j * A[..., 2]).T.astype(np.complex128)      # (ar, ac)
            if ta == "C":
                za = za.conj().T
            zb = (B[..., 0] + 1j * B[..., 2]).T.astype(np.complex128)      # (k, n)
            zc = (C0[..., 0] + 1j * C0[..., 2]).T.astype(np.complex128)
            exact = (2 - 1j) * (za @ zb) + (-3 + 1j) * zc                    # exact below 2^53
            assert np.array_equal(C[..., 0].T, exact.real)
            assert np.array_equal(C[..., 2].T, exact.imag)
            assert np.all(C[..., 1] == 0.0) and np.all(C[..., 3] == 0.0)
    finally:
        oracle.set_order(prev)


def test_complex_k_chunk_flush_period_is_128():
    """The complex AVX2 kernel flushes its accumulators into C every 128 of depth from
    k = 0, as the real one does (wgemm_serial.cpp:207-234): a split of K at 128 or 256 into
    two calls (beta = 1 on the second) is bitwise one call, a split at 64 is not."""
    rng = np.random.default_rng(16)
    m, n, k = 13, 9, 300
    A = ddgen.cplx_matrix(rng, m, k)
    B = ddgen.cplx_matrix(rng, k, n)
    C0 = ddgen.cplx_matrix(rng, m, n)
    al, be = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)
    one = np.array([1.0, 0.0, 0.0, 0.0])
    prev = oracle.set_order(oracle.ORDER_SIMD)
    try:
        full = C0.copy()
        oracle.port_wgemm("N", "N", m, n, k, al, A, m, B, k, be, full, m)

        def split(at):
            C = C0.copy()
            oracle.port_wgemm("N", "N", m, n, at, al, A, m, B, k, be, C, m)
            A2 = np.ascontiguousarray(A[at:])
            B2 = np.ascontiguousarray(B[:, at:, :])
            oracle.port_wgemm("N", "N", m, n, k - at, al, A2, m, B2, k - at, one, C, m)
            return C

        assert np.array_equal(split(128), full)
        assert np.array_equal(split(256), full)
        assert not np.array_equal(split(64), full)
    finally:
        oracle.set_order(prev)


def test_driver_special_cases():
    rng = np.random.default_rng(9)
    m, n, k = 9, 7, 5
    A, B = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n, ld=m + 3)
    C0[2, 3, 0] = np.nan
    zero, one = np.zeros(2), np.array([1.0, 0.0])
    al = ddgen.real_scalar(rng)
    # beta == 0 writes exact zeros (NaN cleared), alpha == 0 leaves only the pre-pass
    C = C0.copy()
    oracle.port_mgemm("N", "N", m, n, k, zero, A, m, B, k, zero, C, m + 3)
    assert np.all(C[:, :m, :] == 0.0) and np.array_equal(C[:, m:, :], C0[:, m:, :])
    # beta == 1 and k == 0: C untouched bit for bit (NaN stays)
    C = C0.copy()
    oracle.port_mgemm("N", "N", m, n, 0, al, A, m, B, k, one, C, m + 3)
    assert np.array_equal(C, C0, equal_nan=True)
    # m == 0 / n == 0: nothing happens
    C = C0.copy()
    oracle.port_mgemm("N", "N", 0, n, k, al, A, m, B, k, zero, C, m + 3)
    assert np.array_equal(C, C0, equal_nan=True)


def test_vector_and_scalar_kernels_and_thread_counts_agree_bitwise():
    rng = np.random.default_rng(10)
    m, n, k = 70, 515, 260
    A, B, C0 = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n), ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng), ddgen.real_scalar(rng)
    L = oracle.lib()
    outs = []
    for force, threads in [(0, 0), (0, 1), (1, 3)]:
        L.port_force_scalar(force)
        C = C0.copy()
        oracle.port_mgemm("N", "N", m, n, k, al, A, m, B, k, be, C, m, nthreads=threads)
        outs.append(C)
    L.port_force_scalar(0)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
