"""Triangular-output GEMM on the GPU (mgemmtr_ / wgemmtr_): bit-for-bit parity with the CPU
restatement of the reference (oracle port_mgemmtr / port_wgemmtr) through the drop-in C ABI
and the device-pointer entry points.

Bar: every element of C identical, padding rows and the complementary triangle included.
The cases follow the reference fuzz test (fuzz_gemmtr_real_body.fypp: n, k log-uniform in
[1,256], random uplo and trans, fuzzed alpha/beta, padded leading dimensions), then medium
and large shapes in every uplo/op combination."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


def _case(rng, cplx, n, k, uplo, ta, tb, scalars=None):
    if scalars is None:
        alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    else:
        alpha, beta = scalars
    ar, ac = (n, k) if ta == "N" else (k, n)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), n + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(uplo=uplo, ta=ta, tb=tb, n=n, k=k, alpha=alpha, beta=beta,
                A=gen(rng, ar, ac, lda), lda=lda, B=gen(rng, br, bc, ldb), ldb=ldb,
                C0=gen(rng, n, n, ldc), ldc=ldc)


def _args(c):
    return (c["uplo"], c["ta"], c["tb"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
            c["ldb"], c["beta"])


def _check(c, cplx, entry=None):
    Cg, Cp = c["C0"].copy(), c["C0"].copy()
    (entry or (mwblas.wgemmtr if cplx else mwblas.mgemmtr))(*_args(c), Cg, c["ldc"])
    (oracle.port_wgemmtr if cplx else oracle.port_mgemmtr)(*_args(c), Cp, c["ldc"])
    assert np.array_equal(Cg, Cp, equal_nan=True), (c["uplo"], c["ta"], c["tb"], c["n"], c["k"])


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemmtr", "wgemmtr"])
def test_fuzz_bitwise_vs_port(gpu, cplx):
    rng = np.random.default_rng(500 + cplx)
    letters = "NTC"
    for i in range(50):
        n, k = ddgen.rand_int_log(rng, 1, 256), ddgen.rand_int_log(rng, 1, 256)
        uplo = "U" if rng.random() < 0.5 else "L"
        ta, tb = letters[int(rng.integers(0, 3))], letters[int(rng.integers(0, 3))]
        print(f"case {i}: {uplo}{ta}{tb} n={n} k={k}", flush=True)
        _check(_case(rng, cplx, n, k, uplo, ta, tb), cplx)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_real_medium_all_ops_bitwise(gpu, uplo):
    rng = np.random.default_rng(510 + (uplo == "U"))
    for ta in "NT":
        for tb in "NT":
            scal = (ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2))
            print(f"medium {uplo}{ta}{tb}", flush=True)
            _check(_case(rng, False, 515, 300, uplo, ta, tb, scal), False)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_grid_rounds_bitwise(gpu, uplo):
    """The rectangle grid has one workgroup per live tile; when the last, partly filled round
    of resident workgroups holds at most half a round, its tiles run as halves.  With 32 x 64
    tiles (three workgroups per CU), n = 2900 leaves a round more than half full (no halves),
    n = 3100 a small one (halves), n = 700 has fewer tiles than one round (all halves); from
    n ~ 5000 the launcher models 64 x 64 tiles (two per CU) as shorter and takes them
    (n = 5000: a last round under half full, 64 x 32 halves).  Small k keeps the restatement
    quick.  Bitwise against the restatement."""
    rng = np.random.default_rng(530 + (uplo == "U"))
    for n, k, ta, tb in [(2900, 96, "N", "N"), (3100, 96, "T", "N"), (700, 96, "N", "T"),
                         (5000, 24, "T", "T")]:
        scal = (ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2))
        print(f"rounds {uplo}{ta}{tb} n={n} k={k}", flush=True)
        _check(_case(rng, False, n, k, uplo, ta, tb, scal), False)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_cplx_medium_bitwise(gpu, uplo):
    rng = np.random.default_rng(520 + (uplo == "U"))
    for ta, tb in [("N", "N"), ("C", "N"), ("T", "C"), ("N", "T")]:
        scal = (ddgen.cplx_scalar(rng, scale=2), ddgen.cplx_scalar(rng, scale=2))
        print(f"complex medium {uplo}{ta}{tb}", flush=True)
        _check(_case(rng, True, 203, 150, uplo, ta, tb, scal), True)


def test_twins_and_device_entry_agree(gpu):
    import torch
    rng = np.random.default_rng(530)
    c = _case(rng, False, 97, 140, "L", "T", "N",
              (ddgen.real_scalar(rng), ddgen.real_scalar(rng)))
    outs = []
    for f in (mwblas.mgemmtr, mwblas.mgemmtr_64,
              lambda *a: mwblas.gemmtr_serial(False, *a)):
        C = c["C0"].copy()
        f(*_args(c), C, c["ldc"])
        outs.append(C)
    dA, dB = torch.from_numpy(c["A"]).to(gpu), torch.from_numpy(c["B"]).to(gpu)
    dC = torch.from_numpy(c["C0"]).to(gpu)
    mwblas.gemmtr_device(False, "L", "T", "N", c["n"], c["k"], c["alpha"], dA, c["lda"], dB,
                         c["ldb"], c["beta"], dC, c["ldc"],
                         stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    outs.append(dC.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0], equal_nan=True)
    Cp = c["C0"].copy()
    oracle.port_mgemmtr(*_args(c), Cp, c["ldc"])
    assert np.array_equal(outs[0], Cp, equal_nan=True)


def test_special_cases(gpu):
    rng = np.random.default_rng(540)
    n, k = 45, 33
    for uplo in "UL":
        for al, be, kk in [([0, 0], [0, 0], k), ([0, 0], [2.5, 1e-20], k), ([1, 0], [0, 0], 0),
                           ([0, 0], [1, 0], k), ([-1, 0], [1, 0], k), ([1.5, 0], [0, 0], k)]:
            c = _case(rng, False, n, k, uplo, "N", "T",
                      (np.array(al, float), np.array(be, float)))
            c["k"] = kk
            c["C0"][7, 3, 0] = np.nan      # column 7, row 3 (upper)
            c["C0"][3, 7, 0] = np.nan      # column 3, row 7 (lower)
            _check(c, False)


def test_large_device_bitwise(gpu):
    """2048 x 2048 triangle, k = 1024, through device pointers (lower 'N','T' and upper
    'T','N'): the whole C against the restatement."""
    import torch
    rng = np.random.default_rng(550)
    n, k = 2048, 1024
    for uplo, ta, tb in [("L", "N", "T"), ("U", "T", "N")]:
        print(f"large {uplo}{ta}{tb}", flush=True)
        c = _case(rng, False, n, k, uplo, ta, tb,
                  (ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)))
        dA, dB = torch.from_numpy(c["A"]).to(gpu), torch.from_numpy(c["B"]).to(gpu)
        dC = torch.from_numpy(c["C0"]).to(gpu)
        mwblas.gemmtr_device(False, uplo, ta, tb, n, k, c["alpha"], dA, c["lda"], dB, c["ldb"],
                             c["beta"], dC, c["ldc"],
                             stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        Cp = c["C0"].copy()
        oracle.port_mgemmtr(*_args(c), Cp, c["ldc"])
        assert np.array_equal(dC.cpu().numpy(), Cp, equal_nan=True)


def test_rectangle_64x64_tile_matches(gpu):
    """MGEMMTR_TILE=64 selects the 64x64 rectangle tile at every size (unset: chosen per call).  The
    setting is read once per process, so the cases run in a child process: 16 fuzz cases
    plus 700 x 900 in both triangles, bitwise against the restatement."""
    import os
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = textwrap.dedent(f"""
        import sys
        sys.path[:0] = [{root!r}, {os.path.join(root, 'epblas-parallel_amd')!r},
                        {os.path.join(root, 'oracle')!r}]
        import numpy as np
        import ddgen, mwblas, oracle
        rng = np.random.default_rng(1500)
        shapes = [(ddgen.rand_int_log(rng, 1, 256), ddgen.rand_int_log(rng, 1, 256)) for _ in range(16)]
        shapes += [(700, 900), (700, 900)]
        for i, (n, k) in enumerate(shapes):
            uplo = "UL"[i % 2]
            ta, tb = "NT"[int(rng.integers(0, 2))], "NT"[int(rng.integers(0, 2))]
            ar, ac = (n, k) if ta == "N" else (k, n)
            br, bc = (k, n) if tb == "N" else (n, k)
            A, B = ddgen.real_matrix(rng, ar, ac, ar + 1), ddgen.real_matrix(rng, br, bc, br + 2)
            C0 = ddgen.real_matrix(rng, n, n, n + 3)
            al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
            args = (uplo, ta, tb, n, k, al, A, ar + 1, B, br + 2, be)
            Cg, Cp = C0.copy(), C0.copy()
            mwblas.mgemmtr(*args, Cg, n + 3)
            oracle.port_mgemmtr(*args, Cp, n + 3)
            assert np.array_equal(Cg, Cp, equal_nan=True), (i, uplo, ta, tb, n, k)
        print("rectangle 64x64 ok", flush=True)
    """)
    env = dict(os.environ, MGEMMTR_TILE="64")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and "rectangle 64x64 ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("nb", [1, 5, 33, 64, 300])
def test_block_sizes_bitwise(gpu, nb):
    """MGEMMTR_NB / WGEMMTR_NB set the reference's diagonal-block size, which fixes which
    entries take the scalar sequence and which the GEMM's; the live-tile grid follows it
    (nb below, equal to and above the 32- and 64-row tiles, and nb > n, where every entry is
    diagonal).  Read once per process, so each nb runs in a child process: real n = 97, 515
    and 2100 (whose last round runs as halves) in both triangles and a complex n = 203,
    bitwise against the restatement with the same nb."""
    import os
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = textwrap.dedent(f"""
        import sys
        sys.path[:0] = [{root!r}, {os.path.join(root, 'epblas-parallel_amd')!r},
                        {os.path.join(root, 'oracle')!r}]
        import numpy as np
        import ddgen, mwblas, oracle
        nb = {nb}
        rng = np.random.default_rng(1600 + nb)
        cases = [(False, 97, 40), (False, 515, 40), (False, 2100, 24), (True, 203, 30)]
        for i, (cplx, n, k) in enumerate(cases):
            for uplo in "UL":
                ta, tb = ("C", "N") if cplx else ("NT"[i % 2], "N")
                ar, ac = (n, k) if ta == "N" else (k, n)
                br, bc = (k, n) if tb == "N" else (n, k)
                gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
                scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
                A, B = gen(rng, ar, ac, ar + 1), gen(rng, br, bc, br + 2)
                C0 = gen(rng, n, n, n + 3)
                al, be = scal(rng, scale=2), scal(rng, scale=2)
                args = (uplo, ta, tb, n, k, al, A, ar + 1, B, br + 2, be)
                Cg, Cp = C0.copy(), C0.copy()
                (mwblas.wgemmtr if cplx else mwblas.mgemmtr)(*args, Cg, n + 3)
                (oracle.port_wgemmtr if cplx else oracle.port_mgemmtr)(*args, Cp, n + 3, nb=nb)
                assert np.array_equal(Cg, Cp, equal_nan=True), (nb, cplx, uplo, n, k)
        print("block size ok", flush=True)
    """)
    env = dict(os.environ, MGEMMTR_NB=str(nb), WGEMMTR_NB=str(nb))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and "block size ok" in r.stdout, r.stdout + r.stderr
