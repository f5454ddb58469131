"""Matrix-vector products on the GPU (mgemv_ / wgemv_): bit-for-bit parity with the CPU
restatement of the reference (oracle port_mgemv / port_wgemv) through the drop-in C ABI and
the device-pointer entry points.

Bar: every element of the y span identical (entries between strided ones included).  The
fuzz cases follow the reference test (fuzz_gemv_real_body.fypp: m, n log-uniform in
[1,256], trans N/T/C, strides 1, 2, 3, -1, -2, fuzzed alpha/beta, padded lda); the real
op-'N' product is also checked for several row-band counts (the reference's OpenMP team
sizes) and at 8192 x 8192."""
import os

import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu


def _omp_default_team():
    """The team an OpenMP runtime forms by default (omp_get_max_threads): the first entry of
    OMP_NUM_THREADS, else the CPUs of the process's affinity mask, capped by
    OMP_THREAD_LIMIT."""
    team = 0
    v = os.environ.get("OMP_NUM_THREADS")
    if v:
        try:
            team = int(v.split(",")[0])
        except ValueError:
            team = 0
    if team <= 0:
        team = len(os.sched_getaffinity(0))
    lim = os.environ.get("OMP_THREAD_LIMIT")
    if lim and lim.isdigit() and 0 < int(lim) < team:
        team = int(lim)
    return team


def _dropin_bands(m):
    """The band count the drop-in uses: MWGEMV_THREADS, else the OpenMP default team."""
    v = os.environ.get("MWGEMV_THREADS")
    team = int(v) if v and v.isdigit() and int(v) > 0 else _omp_default_team()
    return 1 if m < 64 else min(team, 256)


def _vec(rng, length, inc, cplx):
    alloc = 1 + (length - 1) * abs(inc)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return gen(rng, alloc, 1)[0]


def _case(rng, cplx, m=None, n=None, trans=None, incx=None, incy=None):
    m = ddgen.rand_int_log(rng, 1, 256) if m is None else m
    n = ddgen.rand_int_log(rng, 1, 256) if n is None else n
    trans = "NTC"[int(rng.integers(0, 3))] if trans is None else trans
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    lenx, leny = (n, m) if trans == "N" else (m, n)
    lda = m + ddgen.rand_pad(rng)
    incx = ddgen.rand_incx(rng) if incx is None else incx
    incy = ddgen.rand_incx(rng) if incy is None else incy
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(trans=trans, m=m, n=n, alpha=alpha, beta=beta, A=gen(rng, m, n, lda), lda=lda,
                x=_vec(rng, lenx, incx, cplx), incx=incx, y0=_vec(rng, leny, incy, cplx),
                incy=incy)


def _port(c, cplx, nbands):
    y = c["y0"].copy()
    if cplx:
        oracle.port_wgemv(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], c["x"],
                          c["incx"], c["beta"], y, c["incy"])
    else:
        oracle.port_mgemv(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], c["x"],
                          c["incx"], c["beta"], y, c["incy"], nbands)
    return y


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemv", "wgemv"])
def test_fuzz_dropin_bitwise_vs_port(gpu, cplx):
    rng = np.random.default_rng(700 + cplx)
    entry = mwblas.wgemv if cplx else mwblas.mgemv
    for i in range(60):
        c = _case(rng, cplx)
        print(f"case {i}: {c['trans']} m={c['m']} n={c['n']} inc={c['incx']},{c['incy']}",
              flush=True)
        y = c["y0"].copy()
        entry(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], c["x"], c["incx"],
              c["beta"], y, c["incy"])
        expect = _port(c, cplx, _dropin_bands(c["m"]) if c["trans"] == "N" else 1)
        assert np.array_equal(y, expect, equal_nan=True), f"case {i}"


def test_row_bands_on_device(gpu):
    import torch
    rng = np.random.default_rng(710)
    for nb in (1, 3, 16, 100):
        c = _case(rng, False, m=203, n=150, trans="N", incx=1, incy=1)
        dA, dx = torch.from_numpy(c["A"]).to(gpu), torch.from_numpy(c["x"]).to(gpu)
        dy = torch.from_numpy(c["y0"]).to(gpu)
        mwblas.gemv_device(False, "N", c["m"], c["n"], c["alpha"], dA, c["lda"], dx, 1,
                           c["beta"], dy, 1, nbands=nb,
                           stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(dy.cpu().numpy(), _port(c, False, nb), equal_nan=True), nb


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemv", "wgemv"])
def test_large_device_bitwise(gpu, cplx):
    import torch
    rng = np.random.default_rng(720 + cplx)
    m = n = 4096 if cplx else 8192
    for trans in ("N", "C" if cplx else "T"):
        print(f"large {trans} {m}x{n}", flush=True)
        c = _case(rng, cplx, m=m, n=n, trans=trans, incx=1, incy=1)
        scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
        c["alpha"], c["beta"] = scal(rng, scale=2), scal(rng, scale=2)
        dA, dx = torch.from_numpy(c["A"]).to(gpu), torch.from_numpy(c["x"]).to(gpu)
        dy = torch.from_numpy(c["y0"]).to(gpu)
        nb = 16 if trans == "N" else 0
        mwblas.gemv_device(cplx, trans, m, n, c["alpha"], dA, c["lda"], dx, 1, c["beta"], dy, 1,
                           nbands=nb, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(dy.cpu().numpy(), _port(c, cplx, 16), equal_nan=True)


def test_special_cases(gpu):
    rng = np.random.default_rng(730)
    for al, be in [([0, 0], [0, 0]), ([0, 0], [2.5, 1e-20]), ([1, 0], [0, 0]), ([0, 0], [1, 0]),
                   ([-1, 0], [1, 0])]:
        for trans in "NT":
            c = _case(rng, False, m=70, n=33, trans=trans, incx=-2, incy=3)
            c["alpha"], c["beta"] = np.array(al, float), np.array(be, float)
            c["y0"][3, 0] = np.nan
            c["x"][4, 0] = 0.0            # an x entry of zero (skipped column for op 'N')
            y = c["y0"].copy()
            mwblas.mgemv(trans, c["m"], c["n"], c["alpha"], c["A"], c["lda"], c["x"], c["incx"],
                         c["beta"], y, c["incy"])
            nb = _dropin_bands(c["m"]) if trans == "N" else 1
            assert np.array_equal(y, _port(c, False, nb), equal_nan=True), (al, be, trans)
    # m == 0 / n == 0: y untouched, beta not applied
    c = _case(rng, False, m=5, n=4, trans="N", incx=1, incy=1)
    y = c["y0"].copy()
    mwblas.mgemv("N", 5, 0, [1, 0], c["A"], c["lda"], c["x"], 1, [2, 0], y, 1)
    assert np.array_equal(y, c["y0"], equal_nan=True)


def test_twins_agree(gpu):
    import torch
    rng = np.random.default_rng(740)
    c = _case(rng, True, m=99, n=130, trans="C", incx=2, incy=-1)
    outs = []
    for f in (mwblas.wgemv, mwblas.wgemv_64):
        y = c["y0"].copy()
        f(c["trans"], c["m"], c["n"], c["alpha"], c["A"], c["lda"], c["x"], c["incx"],
          c["beta"], y, c["incy"])
        outs.append(y)
    dA, dx = torch.from_numpy(c["A"]).to(gpu), torch.from_numpy(c["x"]).to(gpu)
    dy = torch.from_numpy(c["y0"]).to(gpu)
    mwblas.gemv_device(True, "C", c["m"], c["n"], c["alpha"], dA, c["lda"], dx, 2, c["beta"], dy,
                       -1, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    outs.append(dy.cpu().numpy())
    for o in outs:
        assert np.array_equal(o, _port(c, True, 1), equal_nan=True)


_CHILD = r"""
import os, sys
import numpy as np
os.sched_setaffinity(0, set(sorted(os.sched_getaffinity(0))[:3]))
sys.path[:0] = sys.argv[1:3]
import ddgen, mwblas, oracle
rng = np.random.default_rng(777)
m, n = 203, 150
A = ddgen.real_matrix(rng, m, n)
x = ddgen.real_matrix(rng, n, 1)[0]
y0 = ddgen.real_matrix(rng, m, 1)[0]
al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
y = y0.copy()
mwblas.mgemv("N", m, n, al, A, m, x, 1, be, y, 1)
want = y0.copy()
oracle.port_mgemv("N", m, n, al, A, m, x, 1, be, want, 1, nbands=3)
other = y0.copy()
oracle.port_mgemv("N", m, n, al, A, m, x, 1, be, other, 1, nbands=4)
assert not np.array_equal(want, other), "band count 3 and 4 give the same bits"
assert np.array_equal(y, want), "drop-in band count differs from the 3-CPU affinity mask"
print("affinity child ok")
"""


def test_dropin_bands_follow_a_narrowed_affinity_mask(gpu):
    """A process restricted to three CPUs (taskset, a cgroup cpuset) gets a three-thread
    OpenMP team from the reference, so the drop-in must split op 'N' rows into three bands.
    Runs in a child whose mask is narrowed before the library or the GPU is touched."""
    import subprocess
    import sys
    if len(os.sched_getaffinity(0)) < 4:
        pytest.skip("needs at least four CPUs to narrow the mask to three")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ)
    for name in ("MWGEMV_THREADS", "OMP_NUM_THREADS", "OMP_THREAD_LIMIT"):
        env.pop(name, None)
    print("child: mgemv_ 'N' 203 x 150 under a 3-CPU affinity mask", flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, os.path.join(root, "epblas-parallel_amd"),
                        os.path.join(root, "oracle")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "affinity child ok" in r.stdout
