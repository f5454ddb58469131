"""Pin the CPU restatement of the reference matrix-vector products mgemv_/wgemv_
(oracle/ddgemm_port.c, port_mgemv / port_wgemv).

What pins what:
* accuracy: the reference fuzz tolerance max(10, 4*max(m,n))*2^-104 normwise against the
  binary128 oracle (fuzz_gemv_real_body.fypp: m, n log-uniform in [1,256], random trans
  including 'C', fuzzed alpha/beta, strides 1, 2, 3, -1, -2, padded lda), the product being
  evaluated as a one-column binary128 GEMM on the gathered vectors;
* strides: a strided call equals the unit-stride call on the gathered vectors bit for bit
  (the reference gathers, runs the unit-stride core and scatters, mf_kernels.h:84-93), and
  elements between strided entries are never written;
* exactness: integer operands give the exact integer matrix-vector product;
* op 'N' row bands (mgemv.cpp:60-105): a different team size can change only the rows
  that move between a band's 4-blocks and its tail;
* driver cases: m == 0 or n == 0 leaves y untouched (beta is not applied, mgemv.cpp:250),
  alpha == 0 only scales y, beta == 0 clears NaN.
Bitwise parity of the faithful tail operations with the third-party multifloats operators
themselves is unpinned, as for the GEMMs.
"""
import numpy as np
import pytest

import ddgen
import oracle

TOL_EPS = 2.0 ** -104


def _vec(rng, length, inc, cplx):
    alloc = 1 + (length - 1) * abs(inc)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return gen(rng, alloc, 1)[0]          # (alloc, 2|4)


def _gather(v, length, inc):
    base = -(length - 1) * inc if inc < 0 else 0
    return np.ascontiguousarray(v[[base + i * inc for i in range(length)]])


def _case(rng, cplx):
    m, n = ddgen.rand_int_log(rng, 1, 256), ddgen.rand_int_log(rng, 1, 256)
    trans = "NTC"[int(rng.integers(0, 3))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    lenx, leny = (n, m) if trans == "N" else (m, n)
    lda = m + ddgen.rand_pad(rng)
    incx, incy = ddgen.rand_incx(rng), ddgen.rand_incx(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(trans=trans, m=m, n=n, alpha=alpha, beta=beta, A=gen(rng, m, n, lda), lda=lda,
                x=_vec(rng, lenx, incx, cplx), incx=incx, y0=_vec(rng, leny, incy, cplx),
                incy=incy, lenx=lenx, leny=leny)


def _run(c, cplx, y, nbands=1, x=None, incx=None, incy=None):
    x = c["x"] if x is None else x
    incx = c["incx"] if incx is None else incx
    incy = c["incy"] if incy is None else incy
    if cplx:
        oracle.port_wgemv(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], x, incx,
                          c["beta"], y, incy)
    else:
        oracle.port_mgemv(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], x, incx,
                          c["beta"], y, incy, nbands)


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemv", "wgemv"])
def test_fuzz_against_binary128(cplx):
    rng = np.random.default_rng(600 + cplx)
    for _ in range(60):
        c = _case(rng, cplx)
        y = c["y0"].copy()
        _run(c, cplx, y, nbands=int(rng.integers(1, 17)))
        xg = _gather(c["x"], c["lenx"], c["incx"])
        y0g = _gather(c["y0"], c["leny"], c["incy"])
        yg = _gather(y, c["leny"], c["incy"])
        ta = c["trans"] if cplx else ("N" if c["trans"] == "N" else "T")
        mm, kk = (c["m"], c["n"]) if c["trans"] == "N" else (c["n"], c["m"])
        r = oracle.q128_gemm_err(cplx, ta, "N", mm, 1, kk, c["alpha"], c["A"], c["lda"],
                                 xg[None], kk, c["beta"], y0g[None], mm, yg[None], mm)
        tol = max(10, 4 * max(c["m"], c["n"])) * TOL_EPS
        assert r["err"] <= tol, (c["trans"], c["m"], c["n"], c["incx"], c["incy"], r)


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemv", "wgemv"])
def test_strided_equals_unit_stride_on_gathered_vectors(cplx):
    rng = np.random.default_rng(610 + cplx)
    for _ in range(40):
        c = _case(rng, cplx)
        if c["incx"] == 1 and c["incy"] == 1:
            c["incx"], c["incy"] = -2, 3
            c["x"] = _vec(rng, c["lenx"], -2, cplx)
            c["y0"] = _vec(rng, c["leny"], 3, cplx)
        nb = int(rng.integers(1, 9))
        y = c["y0"].copy()
        _run(c, cplx, y, nbands=nb)
        xg = _gather(c["x"], c["lenx"], c["incx"])
        yg = _gather(c["y0"], c["leny"], c["incy"])
        _run(c, cplx, yg, nbands=nb, x=xg, incx=1, incy=1)
        assert np.array_equal(_gather(y, c["leny"], c["incy"]), yg, equal_nan=True)
        # entries between the strided ones are never written
        touched = np.zeros(len(y), bool)
        base = -(c["leny"] - 1) * c["incy"] if c["incy"] < 0 else 0
        touched[[base + i * c["incy"] for i in range(c["leny"])]] = True
        assert np.array_equal(y[~touched], c["y0"][~touched], equal_nan=True)


def test_integer_operands_give_the_exact_product():
    rng = np.random.default_rng(620)
    for trans, m, n in [("N", 70, 333), ("T", 257, 9), ("C", 5, 64), ("N", 3, 1)]:
        A = ddgen.real_matrix(rng, m, n, kind="integer", bound=64)
        lenx, leny = (n, m) if trans == "N" else (m, n)
        x = ddgen.real_matrix(rng, lenx, 1, kind="integer", bound=64)[0]
        y0 = ddgen.real_matrix(rng, leny, 1, kind="integer", bound=1000)[0]
        y = y0.copy()
        oracle.port_mgemv(trans, m, n, np.array([3.0, 0.0]), A, m, x, 1, np.array([-2.0, 0.0]),
                          y, 1, 4)
        ia = A[..., 0].T.astype(np.int64)
        opa = ia if trans == "N" else ia.T
        exact = 3 * (opa @ x[:, 0].astype(np.int64)) - 2 * y0[:, 0].astype(np.int64)
        assert np.array_equal(y[:, 0], exact.astype(np.float64))
        assert np.all(y[:, 1] == 0.0)


def test_row_bands_only_move_rows_between_block_and_tail():
    rng = np.random.default_rng(630)
    m, n = 203, 150
    A = ddgen.real_matrix(rng, m, n)
    x = ddgen.real_matrix(rng, n, 1)[0]
    y0 = ddgen.real_matrix(rng, m, 1)[0]
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)

    def tail_rows(nb):
        rows = set()
        for t in range(nb):
            lo = t * (m // nb) + min(t, m % nb)
            hi = (t + 1) * (m // nb) + min(t + 1, m % nb)
            rows |= set(range(lo + (hi - lo) // 4 * 4, hi))
        return rows

    ys = {}
    for nb in (1, 3, 7, 16):
        y = y0.copy()
        oracle.port_mgemv("N", m, n, al, A, m, x, 1, be, y, 1, nb)
        ys[nb] = y
    for nb in (3, 7, 16):
        moved = tail_rows(1) ^ tail_rows(nb)
        differ = {i for i in range(m) if not np.array_equal(ys[1][i], ys[nb][i])}
        assert differ <= moved
    # the tail really takes the other (faithful) operations: some moved row differs
    assert any(not np.array_equal(ys[1][i], ys[16][i]) for i in tail_rows(1) ^ tail_rows(16))


def test_driver_special_cases():
    rng = np.random.default_rng(640)
    m, n = 30, 20
    A = ddgen.real_matrix(rng, m, n)
    x = ddgen.real_matrix(rng, n, 1)[0]
    y0 = ddgen.real_matrix(rng, m, 1)[0]
    y0[4, 0] = np.nan
    # m == 0 / n == 0: y untouched even with beta != 1
    for mm, nn in [(0, n), (m, 0)]:
        y = y0.copy()
        oracle.port_mgemv("N", mm, nn, np.array([1.0, 0]), A, m, x, 1, np.array([2.0, 0]), y, 1)
        assert np.array_equal(y, y0, equal_nan=True)
    # alpha == 0: y := beta*y (faithful); beta == 0 clears the NaN
    y = y0.copy()
    be = np.array([2.5, 1e-20])
    oracle.port_mgemv("N", m, n, np.zeros(2), A, m, x, 1, be, y, 1)
    r = np.array([oracle_mul(be, v) for v in y0])
    assert np.array_equal(y, r, equal_nan=True)
    y = y0.copy()
    oracle.port_mgemv("T", m, n, np.zeros(2), A, m, y0[:m], 1, np.zeros(2), y[:n], 1)
    assert np.all(y[:n] == 0.0)


def oracle_mul(a, b):
    import ctypes
    out = np.zeros(2)
    aa, bb = np.ascontiguousarray(a, float), np.ascontiguousarray(b, float)
    oracle.lib().port_dd_mul_exact(aa.ctypes.data_as(ctypes.c_void_p),
                                   bb.ctypes.data_as(ctypes.c_void_p),
                                   out.ctypes.data_as(ctypes.c_void_p))
    return out
