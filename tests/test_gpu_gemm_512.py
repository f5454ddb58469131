"""The reference's own CPU-sized workload on the GPU: real double-double GEMM, M=N=K=512,
NN, column-major, alpha = (1,0), beta = (0,0), entries hi = 2u-1 (u uniform at 53-bit
resolution), lo = hi * 2^-53 * (u' - 1/2).

Bars, on the whole 512 x 512 result (no sampling):
* plain path (drop-in mgemm_ with host operands, and the device entry point): every entry
  bitwise equal to the CPU restatement of the reference mgemm_ (mgemm_parallel.cpp:44-187);
* plain and exact-product paths: normwise error <= 1e-30 against the binary128 oracle with
  the reference metric max|dC| / max(max|C|, |alpha| K max|A| max|B| + |beta| max|C0|)
  (fuzz_util.fypp:252-284), and within the reference fuzz tolerance max(10, 4K) 2^-104."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

N = 512
ALPHA = np.array([1.0, 0.0])
BETA = np.zeros(2)


def _operands():
    rng = np.random.default_rng(512)
    A = ddgen.real_matrix(rng, N, N)
    B = ddgen.real_matrix(rng, N, N)
    return A, B


def _err(A, B, C):
    r = oracle.q128_gemm_err(False, "N", "N", N, N, N, ALPHA, A, N, B, N, BETA,
                             np.zeros((N, N, 2)), N, C, N)
    return r["err"]


def test_plain_dropin_bitwise_and_full_binary128(gpu):
    A, B = _operands()
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        C = np.full((N, N, 2), np.nan)      # beta = 0: NaN in C must be cleared
        mwblas.mgemm("N", "N", N, N, N, ALPHA, A, N, B, N, BETA, C, N)
    finally:
        mwblas.set_path(prev)
    P = np.full((N, N, 2), np.nan)
    oracle.port_mgemm("N", "N", N, N, N, ALPHA, A, N, B, N, BETA, P, N)
    assert np.array_equal(C, P)
    err = _err(A, B, C)
    print(f"512^3 plain drop-in: bitwise equal to the restatement, err {err:.3e}", flush=True)
    assert err <= 1e-30
    assert err <= max(10, 4 * N) * 2.0 ** -104


def test_device_entry_both_paths_full_binary128(gpu):
    import torch
    A, B = _operands()
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    P = np.zeros((N, N, 2))
    oracle.port_mgemm("N", "N", N, N, N, ALPHA, A, N, B, N, BETA, P, N)
    for path, name in ((mwblas.PATH_PLAIN, "plain"), (mwblas.PATH_OZAKI, "exact-product")):
        dC = torch.zeros((N, N, 2), dtype=torch.float64, device=gpu)
        mwblas.gemm_device(False, "N", "N", N, N, N, ALPHA, dA, N, dB, N, BETA, dC, N, path=path)
        torch.cuda.synchronize()
        C = dC.cpu().numpy()
        if path == mwblas.PATH_PLAIN:
            assert np.array_equal(C, P)
        err = _err(A, B, C)
        print(f"512^3 {name} device entry: err {err:.3e}", flush=True)
        assert err <= 1e-30
