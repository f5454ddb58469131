"""Autotuner bookkeeping that needs no GPU (the decision table and its file), and the
tuning runs themselves on the GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

import mwblas


def test_lookup_empty_and_clear():
    mwblas.autotune_clear()
    assert mwblas.autotune_lookup(mwblas.TUNE_PLAIN_REAL, 100, 100, 100) == -1
    assert mwblas.autotune_lookup(99, 1, 1, 1) == -1


def test_decisions_persist_through_the_tune_file(tmp_path):
    f = tmp_path / "tune.txt"
    # kind 0 (plain real), class ceil(log2) of (300, 300, 300) = (9, 9, 9) -> variant 1
    f.write_text("0 9 9 9 1 0.5\n3 11 8 9 17 2.0\n")
    code = ("import sys; sys.path.insert(0, sys.argv[1]); import mwblas;"
            "print(mwblas.autotune_lookup(0, 300, 257, 512), mwblas.autotune_lookup(0, 600, 300, 300),"
            " mwblas.autotune_lookup(3, 2000, 143, 512))")
    env = dict(os.environ, MWGEMM_TUNE_FILE=str(f))
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "epblas-parallel_amd")
    r = subprocess.run([sys.executable, "-c", code, here], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["1", "-1", "17"]


@pytest.mark.gpu
def test_autotune_single_gemm_records_a_variant(gpu):
    mwblas.autotune_clear()
    v, ms = mwblas.autotune(300, 300, 300, path=mwblas.PATH_PLAIN)
    assert v in (0, 1) and ms > 0
    assert mwblas.autotune_lookup(mwblas.TUNE_PLAIN_REAL, 300, 300, 300) == v
    v2, ms2 = mwblas.autotune(600, 600, 600, path=mwblas.PATH_OZAKI)
    assert v2 in (0, 1, 2, 3) and ms2 > 0     # the four residue-GEMM tile variants
    assert mwblas.autotune_lookup(mwblas.TUNE_OZ_GEMM, 768, 768, 640) == v2


@pytest.mark.gpu
def test_tuned_plain_result_still_bitwise(gpu):
    import ddgen
    import oracle
    rng = np.random.default_rng(41)
    m, n, k = 300, 300, 300
    v, _ = mwblas.autotune(m, n, k, path=mwblas.PATH_PLAIN)
    A, B, C0 = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n), ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng), ddgen.real_scalar(rng)
    g, p = C0.copy(), C0.copy()
    mwblas.mgemm("N", "N", m, n, k, al, A, m, B, k, be, g, m)
    oracle.port_mgemm("N", "N", m, n, k, al, A, m, B, k, be, p, m)
    assert np.array_equal(g, p)


@pytest.mark.gpu
def test_batched_autotune_drives_auto_plans(gpu):
    import torch
    mwblas.autotune_clear()
    rng = np.random.default_rng(42)
    cnt = 150
    dims = np.array([[int(np.exp(rng.uniform(np.log(16), np.log(300)))) for _ in range(3)] for _ in range(cnt)])
    ms, ns, ks = dims[:, 0], dims[:, 1], dims[:, 2]
    choice, t = mwblas.batched_autotune("N", "N", ms, ns, ks)
    assert choice in (0, 16, 17, 18) and t > 0
    A = [torch.zeros((int(k), int(m), 2), dtype=torch.float64, device=gpu) for m, k in zip(ms, ks)]
    B = [torch.zeros((int(n), int(k), 2), dtype=torch.float64, device=gpu) for k, n in zip(ks, ns)]
    C = [torch.zeros((int(n), int(m), 2), dtype=torch.float64, device=gpu) for m, n in zip(ms, ns)]
    one = np.tile([1.0, 0.0], (cnt, 1))
    plan = mwblas.BatchedPlan("N", "N", ms, ns, ks, one, A, ms, B, ks, one, C, ms, path=mwblas.PATH_AUTO)
    assert plan.info()["path"] == choice // 16
    plan.close()
