"""Large host-pointer GEMMs on the plain path go through the column-panel pipeline: A up
first, then per panel its columns of op(B) (rows of B when op(B) is transposed) and of C up,
its computation, and its columns of C down, on three streams.

Bars: bitwise equal to the same call through the device-pointer entry point and to the
single-shot staging (MWGEMM_HOST_PANELS=0).  Cases cover real and complex, op(B) = 'N' and
transposed, padded leading dimensions, beta = 0 and beta != 0, a ragged last panel, and
the real and complex exact-product paths (whose panels share op(A)'s maxima and residues).  C's padding
rows must stay untouched."""
import os

import numpy as np
import pytest

import ddgen
import mwblas

pytestmark = pytest.mark.gpu


def _run(cplx, ta, tb, m, n, k, seed, beta_zero, path=mwblas.PATH_PLAIN, device_estimate=None,
         kind="uniform"):
    import torch
    rng = np.random.default_rng(seed)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + 3, br + 5, m + 2
    A, B = gen(rng, ar, ac, lda, kind=kind), gen(rng, br, bc, ldb, kind=kind)
    C0 = gen(rng, m, n, ldc)
    al = scal(rng, scale=2)
    be = np.zeros(4 if cplx else 2) if beta_zero else scal(rng, scale=2)
    print(f"host panels {'wgemm' if cplx else 'mgemm'} {ta}{tb} m={m} n={n} k={k} "
          f"beta_zero={beta_zero} path={path}", flush=True)
    entry = mwblas.wgemm if cplx else mwblas.mgemm
    args = (ta, tb, m, n, k, al, A, lda, B, ldb, be)
    Cp = C0.copy()
    entry(*args, Cp, ldc)                               # panel pipeline
    prev = os.environ.get("MWGEMM_HOST_PANELS")
    os.environ["MWGEMM_HOST_PANELS"] = "0"
    try:
        Cs = C0.copy()
        entry(*args, Cs, ldc)                           # one staged call
    finally:
        if prev is None:
            del os.environ["MWGEMM_HOST_PANELS"]
        else:
            os.environ["MWGEMM_HOST_PANELS"] = prev
    dA, dB, dC = (torch.from_numpy(x).to("cuda:0") for x in (A, B, C0))
    if device_estimate is not None:
        mwblas.ozaki_set_estimate(device_estimate)
    try:
        mwblas.gemm_device(cplx, ta, tb, m, n, k, al, dA, lda, dB, ldb, be, dC, ldc,
                           path=path, stream=0)
        torch.cuda.synchronize()
    finally:
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)
    Cd = dC.cpu().numpy()
    assert np.array_equal(Cp, Cs, equal_nan=True)
    assert np.array_equal(Cp, Cd, equal_nan=True)
    assert np.array_equal(Cp[:, m:, :], C0[:, m:, :])
    return dA, dB, C0, al, be, lda, ldb, ldc


@pytest.mark.parametrize("ta,tb,beta_zero", [("N", "N", False), ("T", "N", True),
                                              ("N", "T", False), ("T", "T", True)])
def test_real_panels_match_one_call(gpu, ta, tb, beta_zero):
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        _run(False, ta, tb, 600, 1100, 1700, 1400 + 2 * (ta == "T") + (tb == "T"), beta_zero)
    finally:
        mwblas.set_path(prev)


@pytest.mark.parametrize("ta,tb", [("C", "N"), ("N", "C")])
def test_complex_panels_match_one_call(gpu, ta, tb):
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        _run(True, ta, tb, 400, 700, 4000, 1410 + (tb == "C"), False)
    finally:
        mwblas.set_path(prev)


@pytest.mark.parametrize("ta,tb,beta_zero", [("N", "N", False), ("T", "T", True)])
def test_exact_product_panels_share_a_and_match_one_call(gpu, ta, tb, beta_zero):
    """The exact-product panels of one call reuse op(A)'s maxima and residues (a tag unique
    to the call); every panel's entries are computed exactly as in one call."""
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    try:
        _run(False, ta, tb, 600, 1100, 1700, 1420 + (ta == "T"), beta_zero, mwblas.PATH_OZAKI)
        # a complex exact-product call in between overwrites the workspace's op(A) part;
        # the next panelled call recomputes it under its own tag
        rng = np.random.default_rng(1425)
        A, B, C = (ddgen.cplx_matrix(rng, 40, 40) for _ in range(3))
        mwblas.wgemm("N", "N", 40, 40, 40, ddgen.cplx_scalar(rng), A, 40, B, 40,
                     ddgen.cplx_scalar(rng), C, 40)
        _run(False, ta, tb, 600, 1100, 1700, 1430 + (ta == "T"), beta_zero, mwblas.PATH_OZAKI)
    finally:
        mwblas.set_path(prev)


def test_dropin_exact_product_keeps_the_worst_case_count(gpu):
    """The drop-in's exact product keeps the worst-case modulus count (its column panels
    would each choose a count from their own columns, and another count rounds the
    reconstruction differently).  With the data-dependent count on for every product, the
    panelled call and the single staged call give the same bits, those of the device call
    with the count off, on wide-exponent operands for which the device call with the count
    on keeps fewer moduli than the worst case."""
    import torch
    m, n, k = 600, 1100, 1700
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS)
    try:
        dA, dB, C0, al, be, lda, ldb, ldc = _run(
            False, "N", "N", m, n, k, 1460, False, mwblas.PATH_OZAKI,
            device_estimate=mwblas.ESTIMATE_OFF, kind="wide")
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS)
        dC = torch.from_numpy(C0).to("cuda:0")
        mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, lda, dB, ldb, be, dC, ldc,
                           path=mwblas.PATH_OZAKI, stream=0)
        assert mwblas.ozaki_last_moduli() < mwblas.ozaki_config(k)[0]
    finally:
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)
        mwblas.set_path(prev)


def test_complex_exact_product_panels_match_one_call(gpu):
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    try:
        _run(True, "C", "N", 400, 700, 4000, 1440, False, mwblas.PATH_OZAKI)
    finally:
        mwblas.set_path(prev)


def test_copy_threads_under_an_openmp_binding(gpu):
    """With OMP_PROC_BIND/OMP_PLACES set, an OpenMP runtime binds the process's initial
    thread to one place, and threads it creates inherit that mask; the host copy pool then
    runs its threads on the process's allotted CPUs instead.  A child process with the
    binding set (and the OpenMP oracle loaded first, so the binding is in force before the
    pool starts) runs a panelled drop-in call, bitwise against the restatement."""
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = textwrap.dedent(f"""
        import os, sys
        sys.path[:0] = [{root!r}, {os.path.join(root, 'epblas-parallel_amd')!r},
                        {os.path.join(root, 'oracle')!r}]
        import numpy as np
        import ddgen, oracle
        oracle.lib()
        import mwblas
        mwblas.set_path(mwblas.PATH_PLAIN)
        rng = np.random.default_rng(1450)
        m, n, k = 600, 1100, 1700
        A, B, C0 = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n), ddgen.real_matrix(rng, m, n)
        al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
        Cg, Cp = C0.copy(), C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, Cg, m)
        oracle.port_mgemm("N", "N", m, n, k, al, A, m, B, k, be, Cp, m)
        assert np.array_equal(Cg, Cp)
        print("bound ok", os.sched_getaffinity(0).__len__(), flush=True)
    """)
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and "bound ok" in r.stdout, r.stdout + r.stderr
