"""Batched DD GEMMs (one persistent launch / one graph replay for many uneven problems).

Bars: plain batched results are bit-identical to the CPU restatement of the reference for
every problem (the batched kernel runs the single-GEMM tile routine); Ozaki batched results
are within 1e-30 normwise of binary128 for every checked problem; graph replays with eager
work between them reproduce the eager result every time."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


def _problems(rng, count, lo, hi, ta="N", tb="N", special=True):
    probs = []
    for i in range(count):
        m, n, k = (ddgen.rand_int_log(rng, lo, hi) for _ in range(3))
        if special and i % 17 == 3:
            k = 0
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = max(ar, 1) + int(rng.integers(0, 3)), max(br, 1) + int(rng.integers(0, 3)), m + int(rng.integers(0, 3))
        al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
        if special and i % 11 == 5:
            be = np.zeros(2)
        if special and i % 13 == 7:
            be = np.array([1.0, 0.0])
        if special and i % 19 == 9:
            al = np.zeros(2)
        probs.append(dict(m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, alpha=al, beta=be,
                          A=ddgen.real_matrix(rng, ar, max(ac, 1), lda),
                          B=ddgen.real_matrix(rng, br, max(bc, 1), ldb),
                          C0=ddgen.real_matrix(rng, m, n, ldc)))
    return probs


def _device(probs, gpu):
    import torch
    dA = [torch.from_numpy(p["A"]).to(gpu) for p in probs]
    dB = [torch.from_numpy(p["B"]).to(gpu) for p in probs]
    dC = [torch.from_numpy(p["C0"]).to(gpu) for p in probs]
    return dA, dB, dC


def _args(probs, dA, dB, dC):
    g = lambda key: [p[key] for p in probs]
    return (g("m"), g("n"), g("k"), np.array(g("alpha")), dA, g("lda"), dB, g("ldb"),
            np.array(g("beta")), dC, g("ldc"))


def _port(p, ta, tb):
    C = p["C0"].copy()
    oracle.port_mgemm(ta, tb, p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                      p["ldb"], p["beta"], C, p["ldc"])
    return C


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_plain_batched_bitwise(gpu, ta, tb):
    import torch
    rng = np.random.default_rng(31)
    probs = _problems(rng, 80, 1, 200, ta, tb)
    dA, dB, dC = _device(probs, gpu)
    mwblas.mgemm_batched_device(ta, tb, *_args(probs, dA, dB, dC), path=mwblas.PATH_PLAIN)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        assert np.array_equal(dC[i].cpu().numpy(), _port(p, ta, tb)), f"problem {i}"


def test_plain_batched_graph_replays(gpu):
    import torch
    rng = np.random.default_rng(32)
    probs = _problems(rng, 50, 1, 160)
    dA, dB, dC = _device(probs, gpu)
    plan = mwblas.BatchedPlan("N", "N", *_args(probs, dA, dB, dC), path=mwblas.PATH_PLAIN)
    expect = [_port(p, "N", "N") for p in probs]
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        for i, p in enumerate(probs):                 # eager work between replays
            dC[i].copy_(torch.from_numpy(p["C0"]))
        torch.cuda.synchronize()
        plan.run(stream)
        torch.cuda.synchronize()
        for i in range(len(probs)):
            assert np.array_equal(dC[i].cpu().numpy(), expect[i]), f"replay {rep} problem {i}"
    plan.close()


@pytest.mark.parametrize("path", [mwblas.PATH_PLAIN, mwblas.PATH_OZAKI], ids=["plain", "ozaki"])
def test_plan_runs_alternating_streams(gpu, path):
    """Runs of one plan share its work counters and workspace.  Three runs alternating
    between two streams with no host synchronisation (each run accumulates into C, beta is
    per problem) must leave C exactly as three runs on one stream do."""
    import torch
    rng = np.random.default_rng(34 + path)
    probs = _problems(rng, 60, 1, 200)
    dA, dB, dC = _device(probs, gpu)
    plan = mwblas.BatchedPlan("N", "N", *_args(probs, dA, dB, dC), path=path)
    one = torch.cuda.current_stream()
    for _ in range(3):
        plan.run(one.cuda_stream)
    torch.cuda.synchronize()
    expect = [c.cpu().numpy() for c in dC]
    for i, p in enumerate(probs):
        dC[i].copy_(torch.from_numpy(p["C0"]))
    torch.cuda.synchronize()
    streams = (torch.cuda.Stream(), torch.cuda.Stream())   # default priority
    for rep in range(3):
        plan.run(streams[rep % 2].cuda_stream)
    torch.cuda.synchronize()
    for i in range(len(probs)):
        assert np.array_equal(dC[i].cpu().numpy(), expect[i], equal_nan=True), f"problem {i}"
    plan.close()


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["ring128", "tile64", "fused64"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_ozaki_batched_vs_binary128(gpu, ta, tb, variant, monkeypatch):
    """Every operand layout of the residue split and maxima (row- and column-contiguous A
    and B), problems of ragged sizes so that unit spaces are padded to whole waves, through
    each residue-GEMM variant (128 x 128 ring tiles, 64 x 64 register-staged tiles, 64 x 64
    ring tiles with the reconstruction fused into them)."""
    import torch
    monkeypatch.setenv("MWGEMM_BATCH_RESIDUE_VARIANT", str(variant))
    rng = np.random.default_rng(33 + 2 * (ta == "T") + (tb == "T"))
    probs = _problems(rng, 40, 1, 300, ta, tb)
    dA, dB, dC = _device(probs, gpu)
    print(f"batched exact product {ta}{tb}, residue variant {variant}", flush=True)
    plan = mwblas.BatchedPlan(ta, tb, *_args(probs, dA, dB, dC), path=mwblas.PATH_OZAKI)
    assert plan.info()["path"] == mwblas.PATH_OZAKI
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(2):
        for i, p in enumerate(probs):
            dC[i].copy_(torch.from_numpy(p["C0"]))
        torch.cuda.synchronize()
        plan.run(stream)
        torch.cuda.synchronize()
        for i, p in enumerate(probs):
            C = dC[i].cpu().numpy()
            if p["m"] == 0 or p["n"] == 0:
                continue
            r = oracle.q128_gemm_err(False, ta, tb, p["m"], p["n"], p["k"], p["alpha"], p["A"],
                                     p["lda"], p["B"], p["ldb"], p["beta"], p["C0"], p["ldc"], C,
                                     p["ldc"])
            assert r["err"] <= 1e-30, (ta, tb, rep, i, p["m"], p["n"], p["k"], r)
            assert np.array_equal(C[:, p["m"]:, :], p["C0"][:, p["m"]:, :])
    plan.close()


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_ozaki_batched_rows_beyond_one_pass(gpu, ta, tb):
    """A row-contiguous operand's maxima and split run in one kernel, one wave per row, 512
    entries per pass held in registers; deeper rows take several passes and read the row
    again for the split.  k in [513, 1400] (NN: B's rows; TT: A's rows), some rows all zero,
    against binary128."""
    import torch
    rng = np.random.default_rng(41 + (ta == "T"))
    probs = _problems(rng, 12, 1, 64, ta, tb, special=False)
    for i, p in enumerate(probs):
        k = ddgen.rand_int_log(rng, 513, 1400)
        ar, ac = (p["m"], k) if ta == "N" else (k, p["m"])
        br, bc = (k, p["n"]) if tb == "N" else (p["n"], k)
        p["k"], p["lda"], p["ldb"] = k, ar + 1, br + 2
        p["A"] = ddgen.real_matrix(rng, ar, ac, ar + 1)
        p["B"] = ddgen.real_matrix(rng, br, bc, br + 2)
        if i % 3 == 1:                      # an all-zero row of the row-contiguous operand
            if tb == "N":
                p["B"][0, :, :] = 0.0       # B column 0 = row 0 of op(B)^T
            if ta == "T":
                p["A"][0, :, :] = 0.0       # A^T row 0
    dA, dB, dC = _device(probs, gpu)
    print(f"batched exact product, deep rows {ta}{tb}", flush=True)
    plan = mwblas.BatchedPlan(ta, tb, *_args(probs, dA, dB, dC), path=mwblas.PATH_OZAKI)
    plan.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i, p in enumerate(probs):
        C = dC[i].cpu().numpy()
        r = oracle.q128_gemm_err(False, ta, tb, p["m"], p["n"], p["k"], p["alpha"], p["A"],
                                 p["lda"], p["B"], p["ldb"], p["beta"], p["C0"], p["ldc"], C,
                                 p["ldc"])
        assert r["err"] <= 1e-30, (ta, tb, i, p["m"], p["n"], p["k"], r)
    plan.close()


def test_batched_workload_2000_uneven(gpu):
    """The batched workload: 2000 problems, M,N,K log-uniform in [16,512], NN, alpha=beta=1,
    entries uniform with non-zero low limbs.  Plain plan: 12 sampled problems bitwise against
    the restatement; Ozaki plan: the same problems within 1e-30 of binary128."""
    import torch
    rng = np.random.default_rng(2000)
    count = 2000
    dims = [(ddgen.rand_int_log(rng, 16, 512), ddgen.rand_int_log(rng, 16, 512),
             ddgen.rand_int_log(rng, 16, 512)) for _ in range(count)]
    gen = torch.Generator(device=gpu)
    gen.manual_seed(7)

    def dev_dd(rows, cols):
        hi = 2 * torch.rand((cols, rows), generator=gen, device=gpu, dtype=torch.float64) - 1
        lo = hi * 2.0 ** -53 * (torch.rand((cols, rows), generator=gen, device=gpu, dtype=torch.float64) - 0.5)
        return torch.stack([hi, lo], -1).contiguous()

    dA = [dev_dd(m, k) for m, n, k in dims]
    dB = [dev_dd(k, n) for m, n, k in dims]
    dC0 = [dev_dd(m, n) for m, n, k in dims]
    ms, ns, ks = [d[0] for d in dims], [d[1] for d in dims], [d[2] for d in dims]
    one = np.tile([1.0, 0.0], (count, 1))
    sample = sorted(rng.choice(count, 12, replace=False).tolist())
    for path in (mwblas.PATH_PLAIN, mwblas.PATH_OZAKI):
        dC = [c.clone() for c in dC0]
        plan = mwblas.BatchedPlan("N", "N", ms, ns, ks, one, dA, ms, dB, ks, one, dC, ms, path=path)
        plan.run(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for i in sample:
            m, n, k = dims[i]
            A, B, C0 = dA[i].cpu().numpy(), dB[i].cpu().numpy(), dC0[i].cpu().numpy()
            C = dC[i].cpu().numpy()
            if path == mwblas.PATH_PLAIN:
                P = C0.copy()
                oracle.port_mgemm("N", "N", m, n, k, one[0], A, m, B, k, one[0], P, m)
                assert np.array_equal(C, P), f"plain problem {i}"
            else:
                r = oracle.q128_gemm_err(False, "N", "N", m, n, k, one[0], A, m, B, k, one[0], C0,
                                         m, C, m)
                assert r["err"] <= 1e-30, (i, dims[i], r)
        plan.close()


def test_batched_rejects_bad_arguments(gpu):
    import torch
    A = torch.zeros((4, 4, 2), dtype=torch.float64, device=gpu)
    with pytest.raises(mwblas.MwgemmError, match="leading dimension"):
        mwblas.mgemm_batched_device("N", "N", [4], [4], [4], np.array([[1.0, 0.0]]), [A], [2],
                                    [A], [4], np.array([[0.0, 0.0]]), [A], [4])
    with pytest.raises(mwblas.MwgemmError, match="trans"):
        mwblas.mgemm_batched_device("Q", "N", [4], [4], [4], np.array([[1.0, 0.0]]), [A], [4],
                                    [A], [4], np.array([[0.0, 0.0]]), [A], [4])


def test_auto_without_a_tuned_choice_goes_by_work(gpu):
    """Untuned AUTO takes the exact product for a batch of at least 1e9 multiply-adds whose
    problems have a mean size of 96 or more (the 2000-problem workload), plain for a batch of
    small problems.  The plans are only built, never run, so one small device tensor stands
    in for every operand."""
    import torch
    mwblas.autotune_clear()
    dummy = torch.zeros((1, 1, 2), dtype=torch.float64, device=gpu)

    def auto_path(dims):
        cnt = len(dims)
        ms, ns, ks = ([d[i] for d in dims] for i in range(3))
        one = np.tile([1.0, 0.0], (cnt, 1))
        plan = mwblas.BatchedPlan("N", "N", ms, ns, ks, one, [dummy] * cnt, ms, [dummy] * cnt,
                                  ks, one, [dummy] * cnt, ms, path=mwblas.PATH_AUTO)
        try:
            return plan.info()["path"]
        finally:
            plan.close()

    rng = np.random.default_rng(2000)
    big = [tuple(ddgen.rand_int_log(rng, 16, 512) for _ in range(3)) for _ in range(2000)]
    assert auto_path(big) == mwblas.PATH_OZAKI
    small = [tuple(int(rng.integers(16, 48)) for _ in range(3)) for _ in range(200)]
    assert auto_path(small) == mwblas.PATH_PLAIN
