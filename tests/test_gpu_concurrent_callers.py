"""Concurrent callers.  The reference's level-3 routines call mgemm_serial from inside their
OpenMP parallel regions (msyrk_serial.cpp:331-356 and the other trailing updates), so
several host threads may enter the library at once.

* Host-pointer calls: eight threads run drop-in calls on arrays of their own, with small
  single-DMA calls and a large call in column panels mixed in.  Each result must be
  bitwise that of the CPU restatement.
* Device-pointer calls: threads run exact-product calls on streams of their own, sharing
  the per-device workspace.  Each result must equal the same call made alone.

ctypes releases the interpreter lock during a foreign call, so the calls really overlap."""
import threading

import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


def _host_case(seed, big=False):
    rng = np.random.default_rng(seed)
    if big:
        m, n, k = 600, 1100, 1700
    else:
        m, n, k = (int(rng.integers(20, 300)) for _ in range(3))
    ta, tb = "NT"[int(rng.integers(0, 2))], "NT"[int(rng.integers(0, 2))]
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A, B = ddgen.real_matrix(rng, ar, ac, ar + 1), ddgen.real_matrix(rng, br, bc, br + 2)
    C0 = ddgen.real_matrix(rng, m, n, m + 3)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    args = (ta, tb, m, n, k, al, A, ar + 1, B, br + 2, be)
    expect = C0.copy()
    oracle.port_mgemm(*args, expect, m + 3)
    return args, C0, m + 3, expect


def test_host_callers_from_eight_threads(gpu):
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        cases = [[_host_case(1600 + 10 * t + r, big=(t == 0 and r == 0)) for r in range(4)]
                 for t in range(8)]
        errors = []

        def worker(t):
            try:
                for r, (args, C0, ldc, expect) in enumerate(cases[t]):
                    C = C0.copy()
                    entry = mwblas.mgemm_serial if (t + r) % 2 else mwblas.mgemm
                    entry(*args, C, ldc)
                    if not np.array_equal(C, expect, equal_nan=True):
                        errors.append((t, r, "differs from the restatement"))
            except Exception as e:   # reported below, in the test's own thread
                errors.append((t, repr(e)))

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=200)
        assert not any(th.is_alive() for th in threads), "a caller thread did not finish"
        assert not errors, errors
    finally:
        mwblas.set_path(prev)


def test_device_callers_on_their_own_streams(gpu):
    import torch
    rng = np.random.default_rng(1700)
    cases = []
    for i in range(6):
        m, n, k = (int(rng.integers(100, 700)) for _ in range(3))
        A, B, C0 = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n), ddgen.real_matrix(rng, m, n)
        al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
        dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
        cases.append((m, n, k, al, be, dA, dB, torch.from_numpy(C0).to(gpu)))

    def run(case, C, stream):
        m, n, k, al, be, dA, dB, _ = case
        mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, m, dB, k, be, C, m,
                           path=mwblas.PATH_OZAKI, stream=stream)

    # each call alone, one after the other
    alone = []
    for c in cases:
        C = c[7].clone()
        torch.cuda.synchronize()
        run(c, C, 0)
        torch.cuda.synchronize()
        alone.append(C.cpu())
    # the same calls from threads, each on its own stream, three rounds each
    streams = [torch.cuda.Stream() for _ in cases]   # default priority
    outs = [[c[7].clone() for _ in range(3)] for c in cases]
    torch.cuda.synchronize()
    errors = []

    def worker(i):
        try:
            for r in range(3):
                run(cases[i], outs[i][r], streams[i].cuda_stream)
        except Exception as e:
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(cases))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=200)
    assert not any(th.is_alive() for th in threads), "a caller thread did not finish"
    torch.cuda.synchronize()
    assert not errors, errors
    for i in range(len(cases)):
        for r in range(3):
            assert torch.equal(outs[i][r].cpu(), alone[i]), (i, r)
