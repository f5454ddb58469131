"""The 256 x 256 residue GEMMs with one workgroup per (modulus, tile) item instead of the
persistent walk (MWGEMM_RING_PERSIST=0, read once per process, hence a child process): each
product is exact against integer arithmetic at ring depths of 1, 6, 7 and 17 stages, and a
whole exact-product DD GEMM (1100 x 1200 x 5000, device pointers) gives the same bits as the
persistent walk in this process."""
import os
import subprocess
import sys

import numpy as np
import pytest

import ddgen
import mwblas

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np, torch
import ddgen, mwblas
dev = torch.device("cuda:0")
m_all, _ = mwblas.ozaki_moduli(40)
for mod_index, kp in ((0, 64), (3, 384), (7, 448), (28, 1088)):
    rng = np.random.default_rng(500 + kp)
    mp, np_ = 256, 512
    m = int(m_all[mod_index])
    lo, hi = (-128, 127) if m == 256 else (-(m - 1) // 2, (m - 1) // 2)
    A = rng.integers(lo, hi + 1, size=(mp, kp), dtype=np.int64)
    B = rng.integers(lo, hi + 1, size=(np_, kp), dtype=np.int64)
    dA = torch.from_numpy(A.astype(np.int8)).to(dev)
    dB = torch.from_numpy(B.astype(np.int8)).to(dev)
    dU = torch.zeros((np_, mp), dtype=torch.uint8, device=dev)
    print(f"one workgroup per item: mod {mod_index} kp {kp}", flush=True)
    mwblas.testing_residue_gemm(dA, dB, dU, mp, np_, kp, mod_index, 0)
    torch.cuda.synchronize()
    got = mwblas.residue_products(dU.cpu().numpy(), mp, np_).astype(np.int64).T
    assert np.array_equal((got - A @ B.T) % m, np.zeros_like(got)), (mod_index, kp)
rng = np.random.default_rng(77)
mm, nn, kk = 1100, 1200, 5000
A = torch.from_numpy(ddgen.real_matrix(rng, mm, kk, kind="wide")).to(dev)
B = torch.from_numpy(ddgen.real_matrix(rng, kk, nn, kind="wide")).to(dev)
C = torch.zeros((nn, mm, 2), dtype=torch.float64, device=dev)
mwblas.gemm_device(False, "N", "N", mm, nn, kk, [1.0, 0.0], A, mm, B, kk, [0.0, 0.0], C, mm,
                   path=mwblas.PATH_OZAKI)
torch.cuda.synchronize()
np.save(sys.argv[2], C.cpu().numpy())
print("child ok", flush=True)
"""


def test_one_workgroup_per_item_matches_the_persistent_walk(gpu, tmp_path):
    out = str(tmp_path / "c_child.npy")
    env = dict(os.environ, MWGEMM_RING_PERSIST="0")
    r = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "epblas-parallel_amd"), out],
                       env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "child ok" in r.stdout
    import torch
    rng = np.random.default_rng(77)
    mm, nn, kk = 1100, 1200, 5000    # 5 x 5 tiles of 256, 29 moduli: the ring path; 79 stages
    A = torch.from_numpy(ddgen.real_matrix(rng, mm, kk, kind="wide")).to("cuda:0")
    B = torch.from_numpy(ddgen.real_matrix(rng, kk, nn, kind="wide")).to("cuda:0")
    C = torch.zeros((nn, mm, 2), dtype=torch.float64, device="cuda:0")
    mwblas.gemm_device(False, "N", "N", mm, nn, kk, [1.0, 0.0], A, mm, B, kk, [0.0, 0.0], C, mm,
                       path=mwblas.PATH_OZAKI)
    torch.cuda.synchronize()
    assert np.array_equal(np.load(out), C.cpu().numpy())
