"""The reference's own GEMM fuzz protocol, run against the GPU library (fuzz_mgemm.fypp,
fuzz_wgemm.fypp, fuzz_util.fypp): per seed, cases with m, n, k log-uniform in [1, 256]
(fuzz_util.fypp:89-110), trans N/T/C (:142-153), alpha/beta in {0, 1, -1} at 10% each else
U[-1,1] (:160-185), leading-dimension pads 0-4 / 5-32 / 33-128 (:390-401), entries
real64x2(2u-1) with a zero low limb (:216-232) and sentinel-filled padding (:43-47,
234-250).  Seeds follow the reference CI rotation (.github/workflows/ci.yml:50-56); as in
the reference, BLAS_FUZZ_SEED picks one seed and BLAS_FUZZ_CASES the cases per seed
(fuzz_util.fypp:51-75; default 80, test/epblas-parallel/CMakeLists.txt:91-93).

Gate, per case: the plain path bit-for-bit equal to the CPU restatement (padding rows
included), and both paths within the reference tolerance max(10, 4K) * 2^-104 of the
binary128 oracle under the reference's error metric err = max|C - C_ref| /
max(max|C_ref|, |alpha| K max|A| max|B| + |beta| max|C0|) (fuzz_util.fypp:252-284)."""
import os

import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

CI_SEEDS = [1, 7, 42, 1234, 20240617]
SEEDS = [int(os.environ["BLAS_FUZZ_SEED"])] if os.environ.get("BLAS_FUZZ_SEED") else CI_SEEDS
CASES = int(os.environ.get("BLAS_FUZZ_CASES", "80"))
EPS = 2.0 ** -104


def _case(rng, cplx):
    m, n, k = (ddgen.rand_int_log(rng, 1, 256) for _ in range(3))
    ta, tb = "NTC"[int(rng.integers(0, 3))], "NTC"[int(rng.integers(0, 3))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), m + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta,
                A=gen(rng, ar, ac, lda, kind="hi_only"), lda=lda,
                B=gen(rng, br, bc, ldb, kind="hi_only"), ldb=ldb,
                C0=gen(rng, m, n, ldc, kind="hi_only"), ldc=ldc)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("cplx", [False, True], ids=["mgemm", "wgemm"])
def test_reference_fuzz_protocol(gpu, cplx, seed):
    rng = np.random.default_rng(seed)
    entry = mwblas.wgemm if cplx else mwblas.mgemm
    port = oracle.port_wgemm if cplx else oracle.port_mgemm
    worst = {mwblas.PATH_PLAIN: 0.0, mwblas.PATH_OZAKI: 0.0}
    for i in range(CASES):
        c = _case(rng, cplx)
        args = (c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
                c["ldb"], c["beta"])
        Cp = c["C0"].copy()
        port(*args, Cp, c["ldc"])
        tol = max(10, 4 * c["k"]) * EPS
        for path in (mwblas.PATH_PLAIN, mwblas.PATH_OZAKI):
            prev = mwblas.set_path(path)
            try:
                Cg = c["C0"].copy()
                entry(*args, Cg, c["ldc"])
            finally:
                mwblas.set_path(prev)
            if path == mwblas.PATH_PLAIN:
                assert np.array_equal(Cg, Cp, equal_nan=True), (seed, i, "plain bitwise")
            else:   # padding rows untouched on the exact-product path too
                assert np.array_equal(Cg[:, c["m"]:, :], c["C0"][:, c["m"]:, :]), (seed, i)
            r = oracle.q128_gemm_err(cplx, c["ta"], c["tb"], c["m"], c["n"], c["k"], c["alpha"],
                                     c["A"], c["lda"], c["B"], c["ldb"], c["beta"], c["C0"],
                                     c["ldc"], Cg, c["ldc"])
            worst[path] = max(worst[path], r["err"] / tol)
            assert r["err"] <= tol, (seed, i, path, c["ta"], c["tb"], c["m"], c["n"], c["k"], r)
    print(f"seed {seed} {'wgemm' if cplx else 'mgemm'}: worst err/tol plain "
          f"{worst[mwblas.PATH_PLAIN]:.3g}, ozaki {worst[mwblas.PATH_OZAKI]:.3g}", flush=True)
