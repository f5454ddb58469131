"""The plain path in the reference's scalar-kernel order (what it runs on hosts without
AVX2/FMA): bit-for-bit parity with the CPU restatement in the same order, for every routine
the order applies to -- mgemm_/wgemm_ (drop-in and device pointers), the triangular-output
GEMM's rectangle, batched plain plans and mgemv_/wgemv_.

Both the library and the restatement are switched to the scalar order for the duration of
each test (mwgemm_set_plain_order / port_set_order) and restored afterwards."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture
def scalar(gpu):
    prev_lib = mwblas.set_plain_order(mwblas.ORDER_SCALAR)
    prev_port = oracle.set_order(oracle.ORDER_SCALAR)
    try:
        yield gpu
    finally:
        mwblas.set_plain_order(prev_lib)
        oracle.set_order(prev_port)


def _case(rng, cplx, mmax=140, kmax=260, letters="NTC"):
    m, n = ddgen.rand_int_log(rng, 1, mmax), ddgen.rand_int_log(rng, 1, mmax)
    k = ddgen.rand_int_log(rng, 1, kmax)
    ta, tb = letters[int(rng.integers(0, len(letters)))], letters[int(rng.integers(0, len(letters)))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    if rng.random() < 0.5:
        alpha = ddgen.cplx_scalar(rng, scale=2) if cplx else ddgen.real_scalar(rng, scale=2)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), m + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, A=gen(rng, ar, ac, lda),
                lda=lda, B=gen(rng, br, bc, ldb), ldb=ldb, C0=gen(rng, m, n, ldc), ldc=ldc)


def _args(c):
    return (c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
            c["ldb"], c["beta"])


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_fuzz_dropin_bitwise(scalar, cplx):
    rng = np.random.default_rng(1100 + cplx)
    entry = mwblas.wgemm if cplx else mwblas.mgemm
    port = oracle.port_wgemm if cplx else oracle.port_mgemm
    for i in range(40):
        c = _case(rng, cplx)
        print(f"scalar order case {i}: {c['ta']}{c['tb']} m={c['m']} n={c['n']} k={c['k']}",
              flush=True)
        Cg, Cp = c["C0"].copy(), c["C0"].copy()
        entry(*_args(c), Cg, c["ldc"])
        port(*_args(c), Cp, c["ldc"])
        assert np.array_equal(Cg, Cp, equal_nan=True), f"case {i}"


def test_real_device_all_ops(scalar):
    import torch
    rng = np.random.default_rng(1110)
    m, n, k = 389, 263, 530
    for ta in "NT":
        for tb in "NT":
            print(f"scalar order device {ta}{tb}", flush=True)
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            A, B = ddgen.real_matrix(rng, ar, ac, ar + 1), ddgen.real_matrix(rng, br, bc, br + 2)
            C0 = ddgen.real_matrix(rng, m, n, m + 3)
            al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
            dA, dB, dC = (torch.from_numpy(x).to(scalar) for x in (A, B, C0))
            mwblas.gemm_device(False, ta, tb, m, n, k, al, dA, ar + 1, dB, br + 2, be, dC, m + 3,
                               stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            Cp = C0.copy()
            oracle.port_mgemm(ta, tb, m, n, k, al, A, ar + 1, B, br + 2, be, Cp, m + 3)
            assert np.array_equal(dC.cpu().numpy(), Cp, equal_nan=True), (ta, tb)


def test_gemmtr_rectangle_follows_the_order(scalar):
    rng = np.random.default_rng(1120)
    for i in range(16):
        cplx = i % 2 == 1
        n, k = ddgen.rand_int_log(rng, 1, 200), ddgen.rand_int_log(rng, 1, 200)
        uplo = "UL"[int(rng.integers(0, 2))]
        letters = "NTC" if cplx else "NT"
        ta, tb = letters[int(rng.integers(0, len(letters)))], letters[int(rng.integers(0, len(letters)))]
        ar, ac = (n, k) if ta == "N" else (k, n)
        br, bc = (k, n) if tb == "N" else (n, k)
        gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
        A, B, C0 = gen(rng, ar, ac, ar + 1), gen(rng, br, bc, br + 1), gen(rng, n, n, n + 2)
        al, be = ddgen.fuzz_alpha_beta(rng, cplx)
        args = (uplo, ta, tb, n, k, al, A, ar + 1, B, br + 1, be)
        Cg, Cp = C0.copy(), C0.copy()
        (mwblas.wgemmtr if cplx else mwblas.mgemmtr)(*args, Cg, n + 2)
        (oracle.port_wgemmtr if cplx else oracle.port_mgemmtr)(*args, Cp, n + 2)
        assert np.array_equal(Cg, Cp, equal_nan=True), (i, uplo, ta, tb, n, k)


def test_batched_plain_follows_the_order(scalar):
    import torch
    rng = np.random.default_rng(1130)
    probs = []
    for _ in range(40):
        m, n, k = (ddgen.rand_int_log(rng, 1, 150) for _ in range(3))
        probs.append(dict(m=m, n=n, k=k, A=ddgen.real_matrix(rng, m, k), B=ddgen.real_matrix(rng, k, n),
                          C0=ddgen.real_matrix(rng, m, n), alpha=ddgen.real_scalar(rng, scale=2),
                          beta=ddgen.real_scalar(rng, scale=2)))
    dA = [torch.from_numpy(p["A"]).to(scalar) for p in probs]
    dB = [torch.from_numpy(p["B"]).to(scalar) for p in probs]
    dC = [torch.from_numpy(p["C0"]).to(scalar) for p in probs]
    g = lambda key: [p[key] for p in probs]
    mwblas.mgemm_batched_device("N", "N", g("m"), g("n"), g("k"), np.array(g("alpha")), dA,
                                g("m"), dB, g("k"), np.array(g("beta")), dC, g("m"),
                                path=mwblas.PATH_PLAIN)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        Cp = p["C0"].copy()
        oracle.port_mgemm("N", "N", p["m"], p["n"], p["k"], p["alpha"], p["A"], p["m"], p["B"],
                          p["k"], p["beta"], Cp, p["m"])
        assert np.array_equal(dC[i].cpu().numpy(), Cp), f"problem {i}"


def _vec(rng, length, inc, cplx):
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return gen(rng, 1 + (length - 1) * abs(inc), 1)[0]


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemv", "wgemv"])
def test_gemv_fuzz_and_large(scalar, cplx):
    import torch
    rng = np.random.default_rng(1140 + cplx)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    cases = []
    for _ in range(30):
        m, n = ddgen.rand_int_log(rng, 1, 256), ddgen.rand_int_log(rng, 1, 256)
        cases.append((m, n, "NTC"[int(rng.integers(0, 3))], ddgen.rand_incx(rng), ddgen.rand_incx(rng)))
    cases += [(3000, 2100, "N", 1, 1), (2100, 3000, "C" if cplx else "T", 1, 1)]
    for i, (m, n, trans, incx, incy) in enumerate(cases):
        print(f"scalar order gemv case {i}: {trans} m={m} n={n}", flush=True)
        lenx, leny = (n, m) if trans == "N" else (m, n)
        lda = m + 1
        A = gen(rng, m, n, lda)
        x, y0 = _vec(rng, lenx, incx, cplx), _vec(rng, leny, incy, cplx)
        x[min(2, len(x) - 1)] = 0.0          # a zero x entry (skipped column for op 'N')
        al, be = ddgen.fuzz_alpha_beta(rng, cplx)
        yg, yp = y0.copy(), y0.copy()
        (mwblas.wgemv if cplx else mwblas.mgemv)(trans, m, n, al, A, lda, x, incx, be, yg, incy)
        if cplx:
            oracle.port_wgemv(trans, m, n, al, A, lda, x, incx, be, yp, incy)
        else:
            oracle.port_mgemv(trans, m, n, al, A, lda, x, incx, be, yp, incy, 1)
        assert np.array_equal(yg, yp, equal_nan=True), (i, trans, m, n, incx, incy)
