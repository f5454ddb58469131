"""Plain (FP64 TwoProd/TwoSum) GPU path: bit-for-bit parity with the CPU restatement of the
reference mgemm_/wgemm_ (oracle/ddgemm_port.c) through the drop-in C ABI, plus the
binary128 accuracy bar at the 4096^3 workload.

Bar: bitwise equality of every C entry (and untouched padding rows) -- the GPU kernel
executes the reference's per-element operation sequence, so any difference is a bug."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


def _case(rng, cplx, mmax=160, kmax=300):
    m, n = ddgen.rand_int_log(rng, 1, mmax), ddgen.rand_int_log(rng, 1, mmax)
    k = ddgen.rand_int_log(rng, 1, kmax)
    letters = "NTC"
    ta, tb = letters[int(rng.integers(0, 3))], letters[int(rng.integers(0, 3))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    if rng.random() < 0.5:   # also random DD scalars with non-zero low limbs
        alpha = ddgen.cplx_scalar(rng, scale=2) if cplx else ddgen.real_scalar(rng, scale=2)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), m + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, A=gen(rng, ar, ac, lda),
                lda=lda, B=gen(rng, br, bc, ldb), ldb=ldb, C0=gen(rng, m, n, ldc), ldc=ldc)


def _run_both(c, cplx, entry):
    C_gpu, C_cpu = c["C0"].copy(), c["C0"].copy()
    args = (c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
            c["ldb"], c["beta"])
    entry(*args, C_gpu, c["ldc"])
    (oracle.port_wgemm if cplx else oracle.port_mgemm)(*args, C_cpu, c["ldc"])
    return C_gpu, C_cpu


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_fuzz_bitwise_vs_port(gpu, cplx):
    rng = np.random.default_rng(1234 + cplx)
    entry = mwblas.wgemm if cplx else mwblas.mgemm
    for i in range(60):
        c = _case(rng, cplx)
        print(f"case {i}: {c['ta']}{c['tb']} m={c['m']} n={c['n']} k={c['k']}", flush=True)
        g, p = _run_both(c, cplx, entry)
        assert np.array_equal(g, p, equal_nan=True), f"case {i} differs"


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_real_medium_bitwise_all_trans(gpu, ta, tb):
    rng = np.random.default_rng(7)
    m, n, k = 515, 263, 700
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    c = dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=ddgen.real_scalar(rng, scale=2),
             beta=ddgen.real_scalar(rng, scale=2), A=ddgen.real_matrix(rng, ar, ac, ar + 3),
             lda=ar + 3, B=ddgen.real_matrix(rng, br, bc, br + 1), ldb=br + 1,
             C0=ddgen.real_matrix(rng, m, n, m + 5), ldc=m + 5)
    g, p = _run_both(c, False, mwblas.mgemm)
    assert np.array_equal(g, p)


@pytest.mark.parametrize("ta,tb", [("C", "N"), ("N", "C"), ("T", "T"), ("C", "C")])
def test_cplx_medium_bitwise(gpu, ta, tb):
    rng = np.random.default_rng(8)
    m, n, k = 130, 97, 300
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    c = dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=ddgen.cplx_scalar(rng, scale=2),
             beta=ddgen.cplx_scalar(rng, scale=2), A=ddgen.cplx_matrix(rng, ar, ac, ar),
             lda=ar, B=ddgen.cplx_matrix(rng, br, bc, br + 2), ldb=br + 2,
             C0=ddgen.cplx_matrix(rng, m, n, m), ldc=m)
    g, p = _run_both(c, True, mwblas.wgemm)
    assert np.array_equal(g, p)


def test_entry_point_twins_agree(gpu):
    """mgemm_, mgemm_64_, mgemm_serial and the device-pointer entry give identical bits."""
    import torch
    rng = np.random.default_rng(9)
    m, n, k = 77, 66, 129
    A, B = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng), ddgen.real_scalar(rng)
    outs = []
    for f in (mwblas.mgemm, mwblas.mgemm_64, mwblas.mgemm_serial):
        C = C0.copy()
        f("N", "N", m, n, k, al, A, m, B, k, be, C, m)
        outs.append(C)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    dC = torch.from_numpy(C0).to(gpu)
    mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, m, dB, k, be, dC, m,
                       stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    outs.append(dC.cpu().numpy())
    # device tensors passed to the drop-in symbol are used in place
    dC2 = torch.from_numpy(C0).to(gpu)
    mwblas.mgemm("N", "N", m, n, k, al, dA, m, dB, k, be, dC2, m)
    outs.append(dC2.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_beta_alpha_special_cases_on_gpu(gpu):
    rng = np.random.default_rng(10)
    m, n, k = 33, 21, 40
    A, B = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n, ld=m + 2)
    C0[3, 4, 0] = np.nan
    for al, be, kk in [([0, 0], [0, 0], k), ([1, 0], [0, 0], k), ([0, 0], [2.5, 1e-20], k),
                       ([1, 0], [1, 0], 0), ([-1, 0], [-1, 0], k)]:
        al, be = np.array(al, float), np.array(be, float)
        g, p = C0.copy(), C0.copy()
        mwblas.mgemm("N", "N", m, n, kk, al, A, m, B, k, be, g, m + 2)
        oracle.port_mgemm("N", "N", m, n, kk, al, A, m, B, k, be, p, m + 2)
        assert np.array_equal(g, p, equal_nan=True)


def test_4096_all_trans_vs_port_and_binary128(gpu):
    """The plain-GPU workload: M=N=K=4096, all four op combinations, random DD alpha/beta
    with hi ~ U[-2,2], C preloaded from the operand distribution.  Full C bitwise against
    the CPU restatement; 256 sampled entries against binary128 (normwise err <= 1e-30)."""
    import torch
    rng = np.random.default_rng(4096)
    m = n = k = 4096
    A = ddgen.real_matrix(rng, m, k)
    B = ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    ent = (rng.integers(0, m, 256), rng.integers(0, n, 256))
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    for ta, tb in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
        print(f"4096 {ta}{tb}", flush=True)
        dC = torch.from_numpy(C0).to(gpu)
        mwblas.gemm_device(False, ta, tb, m, n, k, al, dA, k, dB, k, be, dC, m)
        torch.cuda.synchronize()
        C = dC.cpu().numpy()
        P = C0.copy()
        oracle.port_mgemm(ta, tb, m, n, k, al, A, k, B, k, be, P, m)
        assert np.array_equal(C, P)
        r = oracle.q128_gemm_err(False, ta, tb, m, n, k, al, A, k, B, k, be, C0, m, C, m,
                                 entries=ent)
        print(f"  err vs binary128 = {r['err']:.3e}", flush=True)
        assert r["err"] <= 1e-30


@pytest.mark.parametrize("path", ["plain", "ozaki"])
def test_host_pipeline_chunks_match_device_path(gpu, path):
    """Host operands travel through double-buffered pinned chunks (32 MiB) copied by the
    host worker pool: operands of several chunks, padded leading dimensions, beta != 0 (C
    uploaded too), and columns larger than a chunk.  The result must equal the
    device-pointer call bit for bit, with C's padding rows untouched."""
    import torch
    p = mwblas.PATH_PLAIN if path == "plain" else mwblas.PATH_OZAKI
    rng = np.random.default_rng(31)
    shapes = [(2100, 1100, 1100)]
    if path == "plain":   # (the exact-product workspace would pad n=2 to 256 columns)
        shapes.append((2_200_000, 2, 3))
    for (m, n, k) in shapes:
        print(f"host pipeline {path} m={m} n={n} k={k}", flush=True)
        lda, ldc = m + 5, m + 3
        A = ddgen.real_matrix(rng, m, k, lda)
        B = ddgen.real_matrix(rng, k, n)
        C0 = ddgen.real_matrix(rng, m, n, ldc)
        al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
        prev = mwblas.set_path(p)
        try:
            Ch = C0.copy()
            mwblas.mgemm("N", "N", m, n, k, al, A, lda, B, k, be, Ch, ldc)
        finally:
            mwblas.set_path(prev)
        dA, dB, dC = (torch.from_numpy(x).to(gpu) for x in (A, B, C0))
        mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, lda, dB, k, be, dC, ldc, path=p,
                           stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        Cd = dC.cpu().numpy()
        assert np.array_equal(Ch, Cd, equal_nan=True)
        assert np.array_equal(Ch[:, m:, :], C0[:, m:, :], equal_nan=True)
