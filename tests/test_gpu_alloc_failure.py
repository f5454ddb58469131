"""What happens when device memory cannot be had.

The reference's parallel driver recomputes through its serial core when its pack buffers
cannot be allocated (mgemm_parallel.cpp:180-186), and that serial core leaves C = beta*C
without a word when its own buffers fail (mgemm_serial.cpp:353-365).  Here:
* an exact-product call whose workspace cannot be allocated computes on the plain path,
  which needs no workspace, so C is the plain path's result bit for bit;
* an exact-product batched plan whose workspace cannot be allocated becomes a plain plan;
* a host-pointer call whose device staging cannot be allocated returns an error code
  (MwgemmError in Python) before any copy or kernel, C is left as it was, and the next call
  runs normally.
The failures are produced with the library's allocation-limit hook (every buffer the
library would allocate for itself above the limit counts as unobtainable), so every test
passes ordinary, correctly sized operands.  One test also exhausts the device for real: a
host-pointer call whose 16 TiB staging copy fails before anything is read or written."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

NOMEM = r"cannot allocate"


@pytest.fixture
def alloc_limit():
    """Sets the library's allocation limit for one test and always lifts it afterwards."""
    yield mwblas.testing_set_alloc_limit
    mwblas.testing_set_alloc_limit(0)


def _small(rng, cplx, m=300, n=200, k=400):
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
    return (m, n, k, scal(rng, scale=2), gen(rng, m, k), gen(rng, k, n), scal(rng, scale=2),
            gen(rng, m, n))


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_device_entry_without_workspace_runs_plain(gpu, cplx, alloc_limit):
    import torch
    rng = np.random.default_rng(1200 + cplx)
    m, n, k, al, A, B, be, C0 = _small(rng, cplx)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    stream = torch.cuda.current_stream().cuda_stream

    def call(path):
        dC = torch.from_numpy(C0).to(gpu)
        mwblas.gemm_device(cplx, "N", "N", m, n, k, al, dA, m, dB, k, be, dC, m, path=path,
                           stream=stream)
        torch.cuda.synchronize()
        return dC.cpu().numpy()

    plain = call(mwblas.PATH_PLAIN)
    exact = call(mwblas.PATH_OZAKI)
    assert not np.array_equal(exact, plain), "the exact product should differ in low bits"
    print(f"exact-product call {m}x{n}x{k} with a 1 MiB allocation limit", flush=True)
    alloc_limit(1 << 20)
    fallback = call(mwblas.PATH_OZAKI)
    assert np.array_equal(fallback, plain)
    assert mwblas.last_error()[0] == 0
    alloc_limit(0)
    assert np.array_equal(call(mwblas.PATH_OZAKI), exact)
    p = C0.copy()
    (oracle.port_wgemm if cplx else oracle.port_mgemm)("N", "N", m, n, k, al, A, m, B, k, be, p, m)
    assert np.array_equal(plain, p)


def test_host_entry_staging_failure(gpu, alloc_limit):
    rng = np.random.default_rng(1210)
    m, n, k, al, A, B, be, C0 = _small(rng, False, 70, 50, 90)
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        C1 = C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C1, m)
        print("host entry 70x50x90 with a 64 KiB allocation limit: expect an error", flush=True)
        alloc_limit(64 << 10)
        C2 = C0.copy()
        with pytest.raises(mwblas.MwgemmError, match=NOMEM):
            mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C2, m)
        assert np.array_equal(C2, C0)
        alloc_limit(0)
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C2, m)
        assert np.array_equal(C2, C1)
    finally:
        mwblas.set_path(prev)


def test_host_entry_real_out_of_memory(gpu):
    """The staging copy of a 2^20 x 2^20 host C takes 16 TiB: its allocation fails before
    the one-element operands passed with it are read or written."""
    rng = np.random.default_rng(1211)
    m, n, k, al, A, B, be, C0 = _small(rng, False, 70, 50, 90)
    huge = 1 << 20
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        C1 = C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C1, m)
        one = np.full((1, 1, 2), 3.0)
        guard = one.copy()
        print(f"host entry, m = n = {huge}, k = 1: expect an error", flush=True)
        with pytest.raises(mwblas.MwgemmError, match=NOMEM):
            mwblas.mgemm_64("T", "N", huge, huge, 1, al, one, 1, one, 1, be, guard, huge)
        assert np.array_equal(guard, one)
        C2 = C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C2, m)
        assert np.array_equal(C2, C1)
    finally:
        mwblas.set_path(prev)


def test_batched_plan_without_workspace_runs_plain(gpu, alloc_limit):
    import torch
    rng = np.random.default_rng(1220)
    dims = [(60, 40, 100), (33, 70, 90), (128, 20, 300)]
    alphas = np.array([ddgen.real_scalar(rng, scale=2) for _ in dims])
    betas = np.array([ddgen.real_scalar(rng, scale=2) for _ in dims])
    host = [(ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n), ddgen.real_matrix(rng, m, n))
            for m, n, k in dims]
    ms, ns, ks = zip(*dims)
    stream = torch.cuda.current_stream().cuda_stream

    def run(path):
        dA = [torch.from_numpy(a).to(gpu) for a, _, _ in host]
        dB = [torch.from_numpy(b).to(gpu) for _, b, _ in host]
        dC = [torch.from_numpy(c).to(gpu) for _, _, c in host]
        plan = mwblas.BatchedPlan("N", "N", ms, ns, ks, alphas, dA, ms, dB, ks, betas, dC, ms,
                                  path=path)
        info = plan.info()
        plan.run(stream)
        torch.cuda.synchronize()
        plan.close()
        return info, [c.cpu().numpy() for c in dC]

    _, plain = run(mwblas.PATH_PLAIN)
    info, exact = run(mwblas.PATH_OZAKI)
    assert info["path"] == mwblas.PATH_OZAKI
    print("batched exact-product plan with an 8 KiB allocation limit", flush=True)
    alloc_limit(8 << 10)
    info, fallback = run(mwblas.PATH_OZAKI)
    assert info["path"] == mwblas.PATH_PLAIN
    for f, p in zip(fallback, plain):
        assert np.array_equal(f, p)
    print("batched plan with a 64-byte allocation limit: expect an error", flush=True)
    alloc_limit(64)
    with pytest.raises(mwblas.MwgemmError, match=NOMEM):
        run(mwblas.PATH_PLAIN)
    alloc_limit(0)
    _, again = run(mwblas.PATH_OZAKI)
    for a, e in zip(again, exact):
        assert np.array_equal(a, e)
