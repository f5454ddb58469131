"""Allocation failure is reported, never a fault and never a silent C = beta*C.  The
reference's serial core leaves C = beta*C without a word when its pack buffers cannot be
allocated (mgemm_serial.cpp:353-365, 387-389) and its parallel core recomputes serially
(mgemm_parallel.cpp:180-186).  Here a call whose device memory cannot be had returns an
error code (MwgemmError in Python) before any copy or kernel of the call is enqueued, C is
left as it was, and the next call runs normally, with the same result as before the failure.

The failing calls name sizes whose memory exceeds the device's: the residue products of an
m = n = 2^20 exact-product GEMM take 29 * 2^40 bytes, the staging copy of a 2^20 x 2^20
host C takes 16 TiB, the workspace of a batched plan over one 2^17 x 2^17 problem takes
over 400 GB.  Each of these allocations comes before the first transfer or kernel of its
call, so the one-element operands passed with them are never read or written."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

HUGE = 1 << 20
NOMEM = r"hipMalloc|cannot allocate"


def _small(rng, cplx):
    m, n, k = 70, 50, 90
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
    return (m, n, k, scal(rng, scale=2), gen(rng, m, k), gen(rng, k, n), scal(rng, scale=2),
            gen(rng, m, n))


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_device_entry_workspace_failure(gpu, cplx):
    import torch
    rng = np.random.default_rng(1200 + cplx)
    m, n, k, al, A, B, be, C0 = _small(rng, cplx)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    stream = torch.cuda.current_stream().cuda_stream

    def small_call():
        dC = torch.from_numpy(C0).to(gpu)
        mwblas.gemm_device(cplx, "N", "N", m, n, k, al, dA, m, dB, k, be, dC, m,
                           path=mwblas.PATH_OZAKI, stream=stream)
        torch.cuda.synchronize()
        return dC.cpu().numpy()

    before = small_call()
    w = 4 if cplx else 2
    one = torch.full((1, 1, w), 3.0, dtype=torch.float64, device=gpu)
    guard = one.clone()
    print(f"device entry, exact-product path, m = n = {HUGE}, k = 64: expect an error", flush=True)
    with pytest.raises(mwblas.MwgemmError, match=NOMEM):
        mwblas.gemm_device(cplx, "T", "N", HUGE, HUGE, 64, al, one, 64, one, 64, be, guard, HUGE,
                           path=mwblas.PATH_OZAKI, stream=stream)
    torch.cuda.synchronize()
    assert torch.equal(guard, one)
    assert mwblas.last_error()[0] == 0
    after = small_call()
    assert np.array_equal(after, before)


def test_host_entry_staging_failure(gpu):
    rng = np.random.default_rng(1210)
    m, n, k, al, A, B, be, C0 = _small(rng, False)
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        C1 = C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C1, m)
        one = np.full((1, 1, 2), 3.0)
        guard = one.copy()
        print(f"host entry, m = n = {HUGE}, k = 1: expect an error", flush=True)
        with pytest.raises(mwblas.MwgemmError, match=NOMEM):
            mwblas.mgemm_64("T", "N", HUGE, HUGE, 1, al, one, 1, one, 1, be, guard, HUGE)
        assert np.array_equal(guard, one)
        C2 = C0.copy()
        mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, C2, m)
        assert np.array_equal(C2, C1)
    finally:
        mwblas.set_path(prev)


def test_batched_plan_failure(gpu):
    import torch
    rng = np.random.default_rng(1220)
    one = torch.full((1, 1, 2), 3.0, dtype=torch.float64, device=gpu)
    big = 1 << 17
    alpha = np.array([[1.0, 0.0]])
    print(f"batched plan, exact-product path, one {big} x {big} x 64 problem: expect an error",
          flush=True)
    with pytest.raises(mwblas.MwgemmError, match=NOMEM):
        mwblas.BatchedPlan("T", "N", [big], [big], [64], alpha, [one], [64], [one], [64], alpha,
                           [one], [big], path=mwblas.PATH_OZAKI)
    assert torch.equal(one, torch.full_like(one, 3.0))
    m, n, k, al, A, B, be, C0 = _small(rng, False)
    dA, dB, dC = (torch.from_numpy(x).to(gpu) for x in (A, B, C0))
    plan = mwblas.BatchedPlan("N", "N", [m], [n], [k], al[None], [dA], [m], [dB], [k], be[None],
                              [dC], [m], path=mwblas.PATH_OZAKI)
    plan.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    plan.close()
    r = oracle.q128_gemm_err(False, "N", "N", m, n, k, al, A, m, B, k, be, C0, m,
                             dC.cpu().numpy(), m)
    assert r["err"] <= 1e-30, r
