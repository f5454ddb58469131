"""Stream ordering of the entry points that take device memory.

A NULL stream is HIP's default (null) stream, as for any HIP launch.  A call made with
torch's default stream (handle 0) right after torch kernels produced its operands, with no
synchronisation in between, must read the finished operands.  The torch kernel that reads
the result next on the same stream must see the finished result.  A drop-in (host-entry)
call given device tensors runs on the library's own stream, which is a blocking stream and
is therefore ordered after the default stream.

Each case compares bit for bit with the same call made between full synchronisations."""
import numpy as np
import pytest

import ddgen
import mwblas

pytestmark = pytest.mark.gpu


def _produce(x0):
    """x0's values, written by a chain of exact elementwise kernels on the default stream
    (scaling by 2 and by 1/2), so that the producer is still running when the next call is
    enqueued."""
    y = x0.clone()
    for _ in range(12):
        y.mul_(2.0).mul_(0.5)
    return y


def _operands(gpu, rng, shapes, cplx=False):
    import torch
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return [torch.from_numpy(gen(rng, r, c)).to(gpu) for r, c in shapes]


def _check(gpu, call, host):
    """call(tensors) on the default stream without synchronisation vs between syncs."""
    import torch
    torch.cuda.synchronize()
    ref = [t.clone() for t in host]
    torch.cuda.synchronize()
    call(ref)
    torch.cuda.synchronize()
    ops = [_produce(t) for t in host]
    call(ops)
    got = [t.clone() for t in ops]          # read by a torch kernel on the default stream
    torch.cuda.synchronize()
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("path", [mwblas.PATH_PLAIN, mwblas.PATH_OZAKI], ids=["plain", "ozaki"])
def test_gemm_device_on_the_default_stream(gpu, path):
    rng = np.random.default_rng(1300 + path)
    n, k = 1536, 1024
    host = _operands(gpu, rng, [(n, k), (k, n), (n, n)])
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)

    def call(t):
        mwblas.gemm_device(False, "N", "N", n, n, k, al, t[0], n, t[1], k, be, t[2], n,
                           path=path, stream=0)
    _check(gpu, call, host)


def test_gemmtr_and_gemv_device_on_the_default_stream(gpu):
    rng = np.random.default_rng(1310)
    n, k = 2048, 1024
    host = _operands(gpu, rng, [(n, k), (k, n), (n, n)])
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)

    def tr(t):
        mwblas.gemmtr_device(False, "U", "N", "N", n, k, al, t[0], n, t[1], k, be, t[2], n,
                             stream=0)
    _check(gpu, tr, host)

    vec = _operands(gpu, rng, [(n, n), (n, 1), (n, 1)])

    def mv(t):
        mwblas.gemv_device(False, "N", n, n, al, t[0], n, t[1], 1, be, t[2], 1, stream=0)
    _check(gpu, mv, vec)


def test_dropin_with_device_tensors(gpu):
    rng = np.random.default_rng(1320)
    n, k = 1024, 768
    host = _operands(gpu, rng, [(n, k), (k, n), (n, n)])
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    prev = mwblas.set_path(mwblas.PATH_PLAIN)
    try:
        def call(t):
            mwblas.mgemm("N", "N", n, n, k, al, t[0], n, t[1], k, be, t[2], n)
        _check(gpu, call, host)
    finally:
        mwblas.set_path(prev)
