"""Benchmark: double-double GEMM at 8192^3 on one MI355X, Ozaki-INT8 path vs plain FP64 path.

One step = one C := alpha*A*B + beta*C (real double-double, NN, column-major, M=N=K=8192)
through the exact-product path (INT8 residue GEMMs on the matrix cores + FP64 split and
reconstruction), inputs resident in HBM:
    hi = +-U[1,2) * 2^e, e ~ U_int[-20, 20];  lo = hi * 2^-53 * U[-1/2, 1/2];  alpha=1, beta=0.
Reported value = effective DD-GFLOP/s = 2*M*N*K / t.  The plain FP64 TwoProd/TwoSum kernel
runs on the same inputs for the ratio.  The roofline entry is the dominant kernel (the INT8
residue GEMM) timed with HIP events on the library's stream; the CPU baseline is the
oracle's restatement of the reference OpenMP mgemm_ timed on this host on a bounded row
sample of the same problem (minimum of 3 runs after a warm-up) and on the whole 512^3
workload (minimum of 40 runs after a warm-up).  The same leg, outside every timed region,
checks the normwise error of every GPU result against the binary128 oracle
("err_vs_binary128": the headline's both paths, the complex leg and the plain 4096^3 GEMM
in all four op pairs on sampled entries, the 512^3 workload on all entries, a sample of the
batched problems).

Further objects on the same line (each with its own workload): "gemm_512" (the 512^3
workload on both GPU paths), "library_int8_gemm" (the
vendor INT8 GEMM at one residue GEMM's shape, a reference point for the roofline),
"batched" (2000 small GEMMs in one captured launch per path), "complex" (4096^3 with
op(A)='C', both paths), "gemmtr" (the triangular-output GEMM, 4096 x 4096 upper triangle,
k = 4096), "plain_orders" (the plain GEMM at 4096^3, random alpha/beta, C preloaded: the
reference's AVX2-kernel order in all four op(A)/op(B) N/T pairs and its scalar-kernel order),
"gemv" (16384 x 16384 matrix-vector products, HBM-bound), "host_boundary" (the
drop-in mgemm_ with host operands, PCIe transfers inside the timing) and "serial_trailing"
(a blocked SYRK's trailing updates through mgemm_serial on host arrays).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--size 8192]
                    [--no-batched] [--no-complex] [--no-gemmtr] [--no-gemv] [--no-host]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _sub in ("epblas-parallel_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, _sub))

METRIC = "DD-GEMM GFLOP/s at 8192³: Ozaki-INT8 vs plain FP64-FMA, rel err ≤1e-30"
INT8_PEAK_TOPS = 5000.0      # MI355X dense INT8 MFMA, 5.0 POP/s
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector


def gen_wide(n_rows, n_cols, gen, device):
    import torch
    shape = (n_cols, n_rows)
    sign = torch.where(torch.rand(shape, generator=gen, device=device, dtype=torch.float64) < 0.5,
                       -1.0, 1.0)
    e = torch.randint(-20, 21, shape, generator=gen, device=device, dtype=torch.int32)
    hi = sign * torch.ldexp(1.0 + torch.rand(shape, generator=gen, device=device,
                                             dtype=torch.float64), e)
    lo = hi * 2.0 ** -53 * (torch.rand(shape, generator=gen, device=device,
                                       dtype=torch.float64) - 0.5)
    return torch.stack([hi, lo], dim=-1).contiguous()


ERR_BAR = 1e-30
ERR_METRIC = ("max|C - C_binary128| / max(max|C_binary128|, |alpha| K max|A| max|B| + "
              "|beta| max|C0|) (the reference fuzz metric, fuzz_util.fypp:252-284)")


def cpu_baseline_leg(ah, bh, n, k, gemm512, checks, target_s=4.0, reps=3, reps_512=40):
    """The leg that uses the oracle, outside every timed GPU region.

    CPU baseline: the oracle's CPU restatement of the reference OpenMP mgemm_ (AVX2 kernel,
    jc split over 256-column panels), timed on this host (minimum over repetitions after a
    warm-up, BASELINE protocol): on the first rows of the 8192^3 problem (same inputs), and
    on the whole 512^3 workload.  Checker: the normwise error of every GPU result collected
    in ``checks`` against the binary128 oracle.  Returns (cpu_baseline, err_vs_binary128)."""
    import numpy as np
    import oracle
    L = oracle.lib()
    threads = L.port_max_threads()
    alpha = np.array([1.0, 0.0])
    beta = np.zeros(2)

    def run(rows):
        Ah = np.ascontiguousarray(ah[:, :rows, :])
        C = np.zeros((n, rows, 2))
        t0 = time.perf_counter()
        oracle.port_mgemm("N", "N", rows, n, k, alpha, Ah, rows, bh, k, beta, C, rows)
        return time.perf_counter() - t0

    probe_rows = 64
    t_probe = run(probe_rows)                     # also the warm-up
    rows = int(min(ah.shape[1], max(64, probe_rows * target_s / max(t_probe, 1e-3))))
    rows = max(64, rows // 64 * 64)
    t = min(run(rows) for _ in range(reps))
    gflops = 2.0 * rows * n * k / t / 1e9
    busy = min(threads, (n + 255) // 256)   # the reference splits N into 256-wide panels
    base = {"value": round(gflops, 3), "unit": "DD-GFLOP/s", "cores": busy, "kind": "port",
            "sample": f"rows 0..{rows} of the {ah.shape[1]}x{n}x{k} NN problem (same inputs): "
                      f"minimum of {reps} runs of {t:.2f} s after a {probe_rows}-row warm-up; "
                      f"OMP threads={threads} (jc split keeps {busy} busy), "
                      f"AVX2={bool(L.port_has_avx2())}"}
    if gemm512 is not None:
        a5, b5, m5 = gemm512["A"], gemm512["B"], gemm512["m"]
        c5 = np.zeros((m5, m5, 2))
        oracle.port_mgemm("N", "N", m5, m5, m5, alpha, a5, m5, b5, m5, beta, c5, m5)
        times = []
        for _ in range(reps_512):
            t0 = time.perf_counter()
            oracle.port_mgemm("N", "N", m5, m5, m5, alpha, a5, m5, b5, m5, beta, c5, m5)
            times.append(time.perf_counter() - t0)
        t5 = min(times)
        busy5 = min(threads, (m5 + 255) // 256)
        base["gemm_512"] = {
            "workload": f"real DD GEMM {m5}^3 NN, alpha=(1,0), beta=(0,0), uniform entries, "
                        f"whole problem",
            "value": round(2.0 * m5 ** 3 / t5 / 1e9, 3), "unit": "DD-GFLOP/s",
            "ms_per_call": round(t5 * 1e3, 3), "cores": busy5,
            "protocol": f"one warm-up call, then the minimum of {reps_512} calls "
                        f"(CLOCK_MONOTONIC wall time); OMP threads={threads}, the jc split "
                        f"keeps {busy5} busy",
            "bitwise_equal_to_gpu_plain": bool(np.array_equal(c5, gemm512["C_plain"]))}
    errs = {"bar": ERR_BAR, "metric": ERR_METRIC}
    ok = True
    for name, c in checks.items():
        r = oracle.q128_gemm_err(c["cplx"], c["ta"], c["tb"], c["m"], c["n"], c["k"],
                                 c["alpha"], c["A"], c["lda"], c["B"], c["ldb"], c["beta"],
                                 c["C0"], c["ldc"], c["C"], c["ldc"], entries=c.get("entries"))
        ent = c.get("entries")
        e = {"err": float(r["err"]),
             "entries": "all" if ent is None else f"{len(ent[0])} sampled",
             "ref_tol": max(10, 4 * c["k"]) * 2.0 ** -104}
        if "note" in c:
            e["note"] = c["note"]
        errs[name] = e
        ok = ok and e["err"] <= ERR_BAR
    for name, group in _groups(checks).items():
        errs[name] = {"err": max(errs[g]["err"] for g in group), "problems": len(group),
                      "entries": errs[group[0]]["entries"] + " per problem"}
        for g in group:
            del errs[g]
    errs["all_within_bar"] = bool(ok)
    return base, errs


def _groups(checks):
    """Checks named 'group/i' are reported as one entry per group (the worst error)."""
    out = {}
    for name in checks:
        if "/" in name:
            out.setdefault(name.split("/")[0], []).append(name)
    return out


def _check(cplx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, C, entries=None,
           note=None):
    import numpy as np
    c = {"cplx": cplx, "ta": ta, "tb": tb, "m": m, "n": n, "k": k,
         "alpha": np.asarray(alpha, dtype=np.float64), "A": A, "lda": lda, "B": B, "ldb": ldb,
         "beta": np.asarray(beta, dtype=np.float64), "C0": C0, "ldc": ldc, "C": C,
         "entries": entries}
    if note:
        c["note"] = note
    return c


def gemm_512_bench(dev, reps=20, seed=512):
    """The reference's CPU-sized workload on the GPU: real DD GEMM 512^3 NN, alpha=(1,0),
    beta=(0,0), uniform entries (the same seeded operands as the GPU test of it), both paths,
    device-resident; the best of ``reps`` timed calls.  Returns (bench object, host data for
    the CPU leg and the checker)."""
    import numpy as np
    import torch
    import ddgen
    import mwblas
    m = 512
    rng = np.random.default_rng(seed)
    A = ddgen.real_matrix(rng, m, m)
    B = ddgen.real_matrix(rng, m, m)
    alpha, beta = np.array([1.0, 0.0]), np.zeros(2)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"real DD GEMM {m}^3 NN, alpha=(1,0), beta=(0,0), uniform entries, "
                       f"device-resident", "dd_flops": 2.0 * m ** 3}
    host = {"A": A, "B": B, "m": m}
    for name, path in (("plain", mwblas.PATH_PLAIN), ("ozaki", mwblas.PATH_OZAKI)):
        dC = torch.zeros((m, m, 2), dtype=torch.float64, device=dev)
        mwblas.gemm_device(False, "N", "N", m, m, m, alpha, dA, m, dB, m, beta, dC, m,
                           path=path, stream=stream)
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            mwblas.gemm_device(False, "N", "N", m, m, m, alpha, dA, m, dB, m, beta, dC, m,
                               path=path, stream=stream)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        out[name] = {"value": round(2.0 * m ** 3 / best / 1e9, 2), "unit": "DD-GFLOP/s",
                     "ms_per_call": round(best * 1e3, 4)}
        host["C_" + name] = dC.cpu().numpy()
    return out, host


def batched_bench(dev, reps=20, count=2000, seed=2000, checks=None, check_every=125):
    """The batched workload: count real DD GEMMs, M,N,K log-uniform integers in [16,512],
    NN, alpha=beta=1, uniform entries with non-zero low limbs; one captured plan per path,
    replayed reps times.  Effective DD-GFLOP/s = sum(2MNK) / t.  For the checker, every
    check_every-th problem is restored and the plan run once more outside the timing; those
    problems' results go into ``checks`` (64 sampled entries each)."""
    import numpy as np
    import torch
    import mwblas
    rng = np.random.default_rng(seed)
    lo, hi = np.log(16), np.log(513)
    dims = np.minimum(512, np.maximum(16, np.exp(rng.uniform(lo, hi, size=(count, 3))).astype(np.int64)))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)

    def dd(rows, cols):
        h = 2 * torch.rand((cols, rows), generator=gen, device=dev, dtype=torch.float64) - 1
        l = h * 2.0 ** -53 * (torch.rand((cols, rows), generator=gen, device=dev,
                                         dtype=torch.float64) - 0.5)
        return torch.stack([h, l], -1).contiguous()

    ms, ns, ks = dims[:, 0], dims[:, 1], dims[:, 2]
    A = [dd(int(m), int(k)) for m, k in zip(ms, ks)]
    B = [dd(int(k), int(n)) for k, n in zip(ks, ns)]
    C = [dd(int(m), int(n)) for m, n in zip(ms, ns)]
    one = np.tile([1.0, 0.0], (count, 1))
    flops = float(np.sum(2.0 * ms * ns * ks))
    sample = list(range(0, count, check_every)) if checks is not None else []
    C0s = {i: C[i].clone() for i in sample}
    crng = np.random.default_rng(seed + 1)
    out = {"workload": f"{count} real DD GEMMs, M,N,K log-uniform in [16,512], NN, alpha=beta=1",
           "dd_flops": flops}
    # the library's batched autotuner times plain vs exact product (128 / 64 residue tiles)
    # for this batch class; the exact-product plan then takes the tuned residue tile and the
    # AUTO plan the tuned path
    choice, tuned_ms = mwblas.batched_autotune("N", "N", ms, ns, ks)
    out["autotune"] = {"path": "ozaki" if choice // 16 else "plain",
                       "residue_tile": ("128 x 128", "64 x 64", "64 x 64, reconstruction fused")[
                           choice % 16] if choice // 16 else None,
                       "best_ms": round(tuned_ms, 3)}
    stream = torch.cuda.current_stream().cuda_stream
    for name, path in (("plain", mwblas.PATH_PLAIN), ("ozaki", mwblas.PATH_OZAKI),
                       ("auto", mwblas.PATH_AUTO)):
        plan = mwblas.BatchedPlan("N", "N", ms, ns, ks, one, A, ms, B, ks, one, C, ms, path=path)
        for _ in range(3):                  # warm-up replays (clock and caches settle)
            plan.run(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.run(stream)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        info = plan.info()
        out[name] = {"value": round(flops / t / 1e9, 2), "unit": "DD-GFLOP/s",
                     "ms_per_launch": round(t * 1e3, 3), "tiles": info["tiles"],
                     "moduli": info["moduli"], "launch": "one hipGraph replay"}
        if name == "auto":
            out[name]["path"] = "ozaki" if info["path"] == mwblas.PATH_OZAKI else "plain"
        if sample and name != "auto":
            for i in sample:
                C[i].copy_(C0s[i])
            plan.run(stream)
            torch.cuda.synchronize()
            for i in sample:
                m, n, k = int(ms[i]), int(ns[i]), int(ks[i])
                ent = (crng.integers(0, m, 64), crng.integers(0, n, 64))
                checks[f"batched_{name}/{i}"] = _check(
                    False, "N", "N", m, n, k, [1.0, 0.0], A[i].cpu().numpy(), m,
                    B[i].cpu().numpy(), k, [1.0, 0.0], C0s[i].cpu().numpy(), m,
                    C[i].cpu().numpy(), entries=ent)
        plan.close()
    return out


def gen_uniform(n_rows, n_cols, gen, device, parts=1):
    """Column-major DD (parts=1) or complex DD (parts=2) entries: each hi = 2u-1, lo =
    hi * 2^-53 * (u'-1/2) (non-overlapping); tensor shape (cols, rows, 2*parts)."""
    import torch
    limbs = []
    for _ in range(parts):
        h = 2 * torch.rand((n_cols, n_rows), generator=gen, device=device, dtype=torch.float64) - 1
        l = h * 2.0 ** -53 * (torch.rand((n_cols, n_rows), generator=gen, device=device,
                                         dtype=torch.float64) - 0.5)
        limbs += [h, l]
    return torch.stack(limbs, dim=-1).contiguous()


def random_scalar(rng, parts):
    """A random DD (parts=1) or complex DD (parts=2) scalar, hi ~ U[-2,2]."""
    out = []
    for _ in range(parts):
        h = float(rng.uniform(-2.0, 2.0))
        out += [h, h * 2.0 ** -53 * float(rng.uniform(-0.5, 0.5))]
    return out


def complex_bench(dev, size=4096, oz_reps=5, plain_reps=1, seed=4096, checks=None):
    """Complex DD GEMM C := alpha*A^H*B + beta*C at size^3 (op(A)='C', op(B)='N'), random
    complex alpha and beta, inputs in HBM, on both paths.  Effective rate = 8MNK / t.  Both
    results go into ``checks`` (128 sampled entries) for the checker."""
    import numpy as np
    import torch
    import mwblas
    m = n = k = size
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = gen_uniform(k, m, gen, dev, parts=2)     # op(A) = A^H: A is k x m
    B = gen_uniform(k, n, gen, dev, parts=2)
    C0 = gen_uniform(m, n, gen, dev, parts=2)
    rng = np.random.default_rng(seed)
    alpha, beta = random_scalar(rng, 2), random_scalar(rng, 2)
    stream = torch.cuda.current_stream().cuda_stream
    flops = 8.0 * m * n * k
    out = {"workload": f"complex DD GEMM {m}x{n}x{k}, op(A)='C', op(B)='N', random complex "
                       f"alpha/beta, uniform entries", "dd_flops": flops,
           "moduli": mwblas.ozaki_config(k, cplx=True)[0]}
    res = {}
    for name, path, reps in (("ozaki", mwblas.PATH_OZAKI, oz_reps),
                             ("plain", mwblas.PATH_PLAIN, plain_reps)):
        C = C0.clone()
        mwblas.gemm_device(True, "C", "N", m, n, k, alpha, A, k, B, k, beta, C, m,
                           path=path, stream=stream)
        torch.cuda.synchronize()
        mwblas.timer_reset()
        mwblas.timer_enable(True)
        t = 0.0
        for _ in range(reps):
            C.copy_(C0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mwblas.gemm_device(True, "C", "N", m, n, k, alpha, A, k, B, k, beta, C, m,
                               path=path, stream=stream)
            torch.cuda.synchronize()
            t += time.perf_counter() - t0
        mwblas.timer_enable(False)
        kt = mwblas.timer_read()
        t /= reps
        entry = {"value": round(flops / t / 1e9, 2), "unit": "DD-GFLOP/s",
                 "ms_per_call": round(t * 1e3, 3),
                 "kernels_ms_per_call": {kn: round(v[0] / reps, 3) for kn, v in kt.items()}}
        if name == "plain":
            p_ms, p_cnt = kt.get("plain_cplx", (0.0, 0))
            if p_cnt:
                # 72 FP64 lane-ops per complex DD multiply-add against 39.3e12 lane-ops/s
                entry["fp64_frac"] = round(72 * m * n * k / (p_ms / p_cnt * 1e-3) / 1e12
                                           / (FP64_PEAK_TFLOPS / 2), 4)
        out[name] = entry
        res[name] = C
    if "ozaki" in res and "plain" in res:
        amax = A[..., 0::2].abs().max().item()
        bmax = B[..., 0::2].abs().max().item()
        cmax = C0[..., 0::2].abs().max().item()
        anorm = float(np.hypot(alpha[0], alpha[2]))
        bnorm = float(np.hypot(beta[0], beta[2]))
        d = res["ozaki"] - res["plain"]
        diff = (d[..., 0::2] + d[..., 1::2]).abs().max().item()
        out["ozaki_vs_plain_normwise"] = diff / (anorm * 2 * k * amax * bmax + bnorm * cmax)
        out["speedup_vs_plain"] = round(out["ozaki"]["value"] / out["plain"]["value"], 2)
    if checks is not None:
        crng = np.random.default_rng(seed + 1)
        ent = (crng.integers(0, m, 128), crng.integers(0, n, 128))
        Ah, Bh, C0h = A.cpu().numpy(), B.cpu().numpy(), C0.cpu().numpy()
        for name, C in res.items():
            checks[f"complex_{name}_{size}"] = _check(True, "C", "N", m, n, k, alpha, Ah, k, Bh,
                                                      k, beta, C0h, m, C.cpu().numpy(),
                                                      entries=ent)
    return out


def gemmtr_bench(dev, n=4096, k=4096, reps=3, seed=41):
    """Triangular-output GEMM (mgemmtr_ semantics, plain path, bit-exact): upper triangle
    of an n x n C, NN, random alpha/beta, inputs in HBM.  Effective rate = n(n+1)k / t
    (2 flops per multiply-add over the n(n+1)/2 triangle entries)."""
    import numpy as np
    import torch
    import mwblas
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = gen_uniform(n, k, gen, dev)
    B = gen_uniform(k, n, gen, dev)
    C0 = gen_uniform(n, n, gen, dev)
    rng = np.random.default_rng(seed)
    alpha, beta = random_scalar(rng, 1), random_scalar(rng, 1)
    stream = torch.cuda.current_stream().cuda_stream
    C = C0.clone()
    mwblas.gemmtr_device(False, "U", "N", "N", n, k, alpha, A, n, B, k, beta, C, n, stream=stream)
    torch.cuda.synchronize()
    mwblas.timer_reset()
    mwblas.timer_enable(True)
    t = 0.0
    for _ in range(reps):
        C.copy_(C0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mwblas.gemmtr_device(False, "U", "N", "N", n, k, alpha, A, n, B, k, beta, C, n,
                             stream=stream)
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
    mwblas.timer_enable(False)
    kt = mwblas.timer_read()
    t /= reps
    macs = n * (n + 1) / 2 * k
    return {"workload": f"mgemmtr 'U','N','N' n={n} k={k}, random alpha/beta, plain path",
            "value": round(2 * macs / t / 1e9, 2), "unit": "DD-GFLOP/s",
            "ms_per_call": round(t * 1e3, 3),
            # 18 FP64 lane-ops per DD multiply-add against 39.3e12 lane-ops/s
            "fp64_frac": round(18 * macs / t / 1e12 / (FP64_PEAK_TFLOPS / 2), 4),
            "kernels_ms_per_call": {kn: round(v[0] / reps, 3) for kn, v in kt.items()}}


def plain_orders_bench(dev, n=4096, reps=2, seed=47, checks=None):
    """The plain real GEMM at n^3 (random alpha/beta with hi ~ U[-2,2], C preloaded, inputs in
    HBM): the reference's AVX2/FMA operation order (18 FP64 lane-ops per DD multiply-add) in
    all four op(A)/op(B) N/T pairs, and its scalar order for hosts without AVX2/FMA (29: a
    faithful multiply and a faithful add per multiply-add) on NN.  fp64_frac = lane-ops / t
    against 39.3e12 lane-ops/s.  For the checker, 64 sampled entries of each SIMD-order
    result go into ``checks``."""
    import numpy as np
    import torch
    import mwblas
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A, B, C0 = gen_uniform(n, n, gen, dev), gen_uniform(n, n, gen, dev), gen_uniform(n, n, gen, dev)
    rng = np.random.default_rng(seed)
    alpha, beta = random_scalar(rng, 1), random_scalar(rng, 1)
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"plain real DD GEMM {n}^3, random alpha/beta, C preloaded"}
    prev = mwblas.get_plain_order()
    host = None
    try:
        runs = [("simd_" + ta + tb, mwblas.ORDER_SIMD, 18, ta, tb) for ta in "NT" for tb in "NT"]
        runs.append(("scalar_NN", mwblas.ORDER_SCALAR, 29, "N", "N"))
        for name, order, ops, ta, tb in runs:
            mwblas.set_plain_order(order)
            C = C0.clone()
            mwblas.gemm_device(False, ta, tb, n, n, n, alpha, A, n, B, n, beta, C, n,
                               path=mwblas.PATH_PLAIN, stream=stream)
            torch.cuda.synchronize()
            if checks is not None and order == mwblas.ORDER_SIMD:
                if host is None:
                    crng = np.random.default_rng(seed + 1)
                    host = (A.cpu().numpy(), B.cpu().numpy(), C0.cpu().numpy(),
                            (crng.integers(0, n, 64), crng.integers(0, n, 64)))
                Ah, Bh, C0h, ent = host
                checks[f"plain_{n}_ops/{ta}{tb}"] = _check(False, ta, tb, n, n, n, alpha, Ah, n,
                                                           Bh, n, beta, C0h, n, C.cpu().numpy(),
                                                           entries=ent)
            t0 = time.perf_counter()
            for _ in range(reps):
                mwblas.gemm_device(False, ta, tb, n, n, n, alpha, A, n, B, n, beta, C, n,
                                   path=mwblas.PATH_PLAIN, stream=stream)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / reps
            out[name] = {"value": round(2.0 * n ** 3 / t / 1e9, 2), "unit": "DD-GFLOP/s",
                         "ms_per_call": round(t * 1e3, 2),
                         "fp64_frac": round(ops * n ** 3 / t / 1e12 / (FP64_PEAK_TFLOPS / 2), 4)}
    finally:
        mwblas.set_plain_order(prev)
    return out


def gemv_bench(dev, n=16384, reps=5, seed=43):
    """Matrix-vector products (mgemv_ semantics, bit-exact, 16 row bands for op 'N'): an
    n x n real DD A in HBM, ops 'N' and 'T'.  HBM-bound: A (16 B per entry) is read once, so
    the rate is reported as A-bytes / t against the 8 TB/s HBM peak."""
    import numpy as np
    import torch
    import mwblas
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A = gen_uniform(n, n, gen, dev)
    x = gen_uniform(n, 1, gen, dev)
    y0 = gen_uniform(n, 1, gen, dev)
    rng = np.random.default_rng(seed)
    alpha, beta = random_scalar(rng, 1), random_scalar(rng, 1)
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"mgemv {n}x{n}, random alpha/beta, unit strides", "bytes_A": 16 * n * n}
    for trans in ("N", "T"):
        y = y0.clone()
        mwblas.gemv_device(False, trans, n, n, alpha, A, n, x, 1, beta, y, 1, nbands=16,
                           stream=stream)
        torch.cuda.synchronize()
        t = 0.0
        for _ in range(reps):
            y.copy_(y0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mwblas.gemv_device(False, trans, n, n, alpha, A, n, x, 1, beta, y, 1, nbands=16,
                               stream=stream)
            torch.cuda.synchronize()
            t += time.perf_counter() - t0
        t /= reps
        out[trans] = {"ms_per_call": round(t * 1e3, 4),
                      "GBps": round(16 * n * n / t / 1e9, 1),
                      "hbm_frac": round(16 * n * n / t / 1e9 / 8000.0, 4),
                      "dd_gflops": round(2.0 * n * n / t / 1e9, 2)}
    return out


def host_boundary_bench(size=4096, reps=3, seed=77):
    """The drop-in mgemm_ with host (pageable) operands: H2D of A and B, the GEMM on the
    exact-product path, D2H of C, all inside the timed call (PCIe-inclusive).  Real NN at
    size^3, alpha=1, beta=0.  Reported beside, never instead of, the HBM-resident value."""
    import numpy as np
    import mwblas
    rng = np.random.default_rng(seed)

    def dd(rows, cols):
        x = np.empty((cols, rows, 2))
        x[..., 0] = rng.uniform(-1.0, 1.0, (cols, rows))
        x[..., 1] = x[..., 0] * 2.0 ** -53 * (rng.random((cols, rows)) - 0.5)
        return x

    m = n = k = size
    A, B = dd(m, k), dd(k, n)
    C = np.zeros((n, m, 2))
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    try:
        mwblas.mgemm("N", "N", m, n, k, [1.0, 0.0], A, m, B, k, [0.0, 0.0], C, m)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            mwblas.mgemm("N", "N", m, n, k, [1.0, 0.0], A, m, B, k, [0.0, 0.0], C, m)
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
    finally:
        mwblas.set_path(prev)
    moved = (A.nbytes + B.nbytes + C.nbytes)
    # The default (plain) path on the same host operands.  Large calls on either path go
    # through the column-panel pipeline, which overlaps the panels' transfers with their
    # computation.
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        mwblas.mgemm("N", "N", m, n, k, [1.0, 0.0], A, m, B, k, [0.0, 0.0], C, m)
        best_plain = None
        for _ in range(max(1, reps - 1)):
            t0 = time.perf_counter()
            mwblas.mgemm("N", "N", m, n, k, [1.0, 0.0], A, m, B, k, [0.0, 0.0], C, m)
            t = time.perf_counter() - t0
            best_plain = t if best_plain is None else min(best_plain, t)
    finally:
        mwblas.set_path(prev)
    return {"workload": f"mgemm_ NN {m}x{n}x{k} with host operands (pageable), Ozaki path, "
                        f"alpha=1, beta=0", "value": round(2.0 * m * n * k / best / 1e9, 2),
            "unit": "DD-GFLOP/s", "ms_per_call": round(best * 1e3, 2),
            "host_bytes_moved": moved,
            "effective_GBps": round(moved / best / 1e9, 2),
            "plain_path": {"value": round(2.0 * m * n * k / best_plain / 1e9, 2),
                           "unit": "DD-GFLOP/s", "ms_per_call": round(best_plain * 1e3, 2),
                           "pipeline": "column panels (transfers overlap computation)"}}


def library_int8_bench(dev, size, reps=20):
    """A plain INT8 x INT8 -> INT32 GEMM of one residue GEMM's shape (size^3) through the
    vendor library (torch._int_mm, hipBLASLt on ROCm), as a reference point for what this
    GPU sustains on INT8 MFMA at this shape: int8 ops per second and the fraction of the
    5.0 POP/s peak.  Nothing of it is used by the library."""
    import torch
    gen = torch.Generator(device=dev)
    gen.manual_seed(size)
    a = torch.randint(-127, 128, (size, size), dtype=torch.int8, device=dev, generator=gen)
    b = torch.randint(-127, 128, (size, size), dtype=torch.int8, device=dev,
                      generator=gen).t().contiguous().t()
    for _ in range(3):
        torch._int_mm(a, b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch._int_mm(a, b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tops = 2.0 * size ** 3 / (ms * 1e-3) / 1e12
    return {"workload": f"torch._int_mm int8 {size}^3 (B column-major), int32 output",
            "ms": round(ms, 3), "TOP/s": round(tops, 1),
            "frac": round(tops / INT8_PEAK_TOPS, 4)}


def serial_trailing_bench(n=4096, k=1024, nb=64, seed=78):
    """The trailing updates of a blocked lower SYRK, C(j0:n, jc:jc+nb) += alpha *
    A(j0:n, :) A(jc:jc+nb, :)^T, issued the way msyrk_serial.cpp:331-339 issues them: one
    mgemm_serial call per 64-column block on interior pointers of host arrays (the
    diagonal-block kernels between them are not part of this library and are not run).
    Every call stages its operands over PCIe.  Also the cost of one tiny 8x8x8 call (the
    median of individually timed calls, measured first: right after a few hundred MB of
    fresh host arrays appear, calls are transiently slower for reasons outside the
    library)."""
    import numpy as np
    import mwblas
    alpha, one = np.array([1.0, 0.0]), np.array([1.0, 0.0])
    tiny = np.zeros((8, 8, 2))
    tiny[..., 0] = 1.0
    out = np.zeros((8, 8, 2))
    times = []
    for i in range(220):
        t1 = time.perf_counter()
        mwblas.mgemm_serial("N", "N", 8, 8, 8, alpha, tiny, 8, tiny, 8, one, out, 8)
        if i >= 20:
            times.append(time.perf_counter() - t1)
    tiny_us = float(np.median(times)) * 1e6
    rng = np.random.default_rng(seed)
    lda = ldc = n
    A = np.empty((k, n, 2))
    A[..., 0] = rng.uniform(-1.0, 1.0, (k, n))
    A[..., 1] = A[..., 0] * 2.0 ** -53 * (rng.random((k, n)) - 0.5)
    C = np.zeros((n, n, 2))
    flat_a, flat_c = A.reshape(-1, 2), C.reshape(-1, 2)
    blocks = [(jc, min(nb, n - jc)) for jc in range(0, n, nb) if n - jc - min(nb, n - jc) > 0]

    def sequence():
        for jc, jb in blocks:
            j0 = jc + jb
            mwblas.mgemm_serial("N", "T", n - j0, jb, k, alpha, flat_a[j0:], lda, flat_a[jc:],
                                lda, one, flat_c[jc * ldc + j0:], ldc)

    sequence()
    t0 = time.perf_counter()
    sequence()
    t = time.perf_counter() - t0
    macs = sum((n - jc - jb) * jb * k for jc, jb in blocks)
    return {"workload": f"lower-SYRK trailing updates n={n} k={k} nb={nb}: {len(blocks)} "
                        f"mgemm_serial('N','T') calls on host arrays",
            "value": round(2.0 * macs / t / 1e9, 2), "unit": "DD-GFLOP/s",
            "ms_total": round(t * 1e3, 2), "calls": len(blocks),
            "us_per_tiny_call_8x8x8": round(tiny_us, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-complex", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-gemmtr", action="store_true")
    ap.add_argument("--no-gemv", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run one device copy of this many MiB (the known-size kernel "
                         "that scales hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the benchmark runs on exactly one GPU; several GPUs would each take a "
              "column block of B and C, with no exchange beyond a broadcast of A. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import mwblas

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.calib_copy_mb > 0:
        src = torch.ones(args.calib_copy_mb << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n = k = m = args.size
    gen = torch.Generator(device=dev)
    gen.manual_seed(8192)
    A = gen_wide(m, k, gen, dev)          # (cols=k, ld=m, 2)
    B = gen_wide(k, n, gen, dev)          # (cols=n, ld=k, 2)
    C = torch.zeros((n, m, 2), dtype=torch.float64, device=dev)
    alpha, beta = [1.0, 0.0], [0.0, 0.0]
    stream = torch.cuda.current_stream().cuda_stream
    nmod, bits = mwblas.ozaki_config(k)
    # the library's autotuner picks the residue-GEMM tile for this shape class
    tile_variant, _, tile_ms = mwblas.autotune(m, n, k, path=mwblas.PATH_OZAKI, all_times=True)

    def step(path, out):
        mwblas.gemm_device(False, "N", "N", m, n, k, alpha, A, m, B, k, beta, out, m,
                           path=path, stream=stream)

    for _ in range(args.warmup):
        step(mwblas.PATH_OZAKI, C)
    torch.cuda.synchronize()
    mwblas.timer_reset()
    mwblas.timer_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(mwblas.PATH_OZAKI, C)
    torch.cuda.synchronize()
    t_oz = (time.perf_counter() - t0) / args.steps
    mwblas.timer_enable(False)
    kt = mwblas.timer_read()
    # moduli kept by the data-dependent count (the residues are formed for nmod; the residue
    # GEMMs and the reconstruction of the others do not run)
    kept = mwblas.ozaki_last_moduli() or nmod

    # plain FP64 path on the same inputs
    kp, t_plain, agree = {}, None, None
    if not args.skip_plain:
        Cp = torch.zeros_like(C)
        step(mwblas.PATH_PLAIN, Cp)
        torch.cuda.synchronize()
        mwblas.timer_reset()
        mwblas.timer_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            step(mwblas.PATH_PLAIN, Cp)
        torch.cuda.synchronize()
        t_plain = (time.perf_counter() - t0) / args.plain_steps
        mwblas.timer_enable(False)
        kp = mwblas.timer_read()
        # agreement of the two GPU results, normwise (|alpha| K max|A| max|B|)
        amax = A[..., 0].abs().max().item()
        bmax = B[..., 0].abs().max().item()
        diff = ((C[..., 0] - Cp[..., 0]) + (C[..., 1] - Cp[..., 1])).abs().max().item()
        agree = diff / (k * amax * bmax)

    # results for the checker (outside every timed region): 256 sampled entries each
    checks = {}
    if not args.no_cpu:
        import numpy as np
        ent = (np.random.default_rng(8193).integers(0, m, 256),
               np.random.default_rng(8194).integers(0, n, 256))
        ah, bh = A.cpu().numpy(), B.cpu().numpy()
        zero = np.zeros((n, m, 2))          # C0 (beta = 0)
        for name, out in (("ozaki", C), ("plain", None if args.skip_plain else Cp)):
            if out is not None:
                checks[f"{name}_{m}"] = _check(False, "N", "N", m, n, k, alpha, ah, m, bh, k,
                                               beta, zero, m, out.cpu().numpy(), entries=ent)

    flops = 2.0 * m * n * k
    oz_gflops = flops / t_oz / 1e9
    plain_gflops = flops / t_plain / 1e9 if t_plain else None
    g_ms, g_cnt = kt.get("oz_gemm", (0.0, 0))
    g_avg = g_ms / max(g_cnt, 1)
    ops = 2.0 * m * n * k * kept                         # INT8 ops per residue-GEMM launch
    achieved = ops / (g_avg * 1e-3) / 1e12 if g_avg > 0 else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "oz_gemm_8192_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    p_ms, p_cnt = kp.get("plain_real", (0.0, 0))

    result = {
        "metric": METRIC,
        "value": round(oz_gflops, 2),
        "unit": "DD-GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_oz * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-mfma+f64",
        "data": "synthetic",
        "config": {"workload": f"ozaki_int8_real_dd_gemm_{m}", "m": m, "n": n, "k": k,
                   "trans": "NN", "alpha": alpha, "beta": beta,
                   "inputs": "hi=+-U[1,2)*2^U_int[-20,20], lo=hi*2^-53*U[-1/2,1/2]",
                   "moduli": kept, "moduli_worst_case": nmod, "bits": bits,
                   "parallelism": "single-gpu",
                   "residue_tile_variant": tile_variant,
                   "autotune_ms_per_variant": [round(x, 3) for x in tile_ms]},
        "plain_fp64": None if not t_plain else {
            "value": round(plain_gflops, 2), "unit": "DD-GFLOP/s",
            "ms_per_step": round(t_plain * 1e3, 2),
            "kernel_ms": round(p_ms / max(p_cnt, 1), 2),
            # 18 FP64 lane-ops per DD multiply-add against the FP64 vector peak
            # (78.6 TFLOP/s counts an FMA as 2, i.e. 39.3e12 lane-ops/s)
            "fp64_frac": round(18 * m * n * k / (p_ms / max(p_cnt, 1) * 1e-3) / 1e12
                               / (FP64_PEAK_TFLOPS / 2), 4) if p_cnt else None},
        "speedup_vs_plain": round(oz_gflops / plain_gflops, 2) if t_plain else None,
        "ozaki_vs_plain_normwise": agree,
        "kernels_ms_per_step": {name: round(v[0] / max(v[1], 1) * (v[1] / args.steps), 3)
                                for name, v in kt.items()},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1) if achieved else None,
                     "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / INT8_PEAK_TOPS, 4) if achieved else None,
                     "traffic": traffic, "kernel": "k_oz_gemm_ring_persist (INT8 residue GEMMs, int8 ops)"},
    }
    try:
        result["library_int8_gemm"] = library_int8_bench(dev, n)
    except (RuntimeError, AttributeError) as e:   # reference point only
        result["library_int8_gemm"] = {"error": str(e)[:200]}
    gemm512 = None
    result["gemm_512"], gemm512 = gemm_512_bench(dev)
    if not args.no_cpu:
        for name in ("plain", "ozaki"):
            checks[f"{name}_512"] = _check(False, "N", "N", 512, 512, 512, [1.0, 0.0],
                                           gemm512["A"], 512, gemm512["B"], 512, [0.0, 0.0],
                                           np.zeros((512, 512, 2)), 512, gemm512["C_" + name])
    if not args.no_batched:
        result["batched"] = batched_bench(dev, checks=None if args.no_cpu else checks)
    if not args.no_complex:
        result["complex"] = complex_bench(dev, checks=None if args.no_cpu else checks)
    if not args.no_gemmtr:
        result["gemmtr"] = gemmtr_bench(dev)
        result["plain_orders"] = plain_orders_bench(dev, checks=None if args.no_cpu else checks)
    if not args.no_gemv:
        result["gemv"] = gemv_bench(dev)
    if not args.no_host:
        result["host_boundary"] = host_boundary_bench()
        result["serial_trailing"] = serial_trailing_bench()
    if not args.no_cpu:
        result["cpu_baseline"], result["err_vs_binary128"] = cpu_baseline_leg(
            ah, bh, n, k, gemm512, checks)
    print(json.dumps(result))


if __name__ == "__main__":
    # No OMP_PROC_BIND here: binding pins the main thread to one core, and the library's
    # host copy threads, created from it, would inherit that one-core mask.
    main()
