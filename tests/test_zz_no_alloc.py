"""The no-allocation entry point mwgemm_mgemm_device_noalloc, called through the C ABI.

Bar: the same result as the existing device entry point mwgemm_mgemm_device on the same
inputs, bit for bit (the exact-product path runs the same kernels, with its workspace in the
library's static scratch memory instead of an allocated one; the plain path needs none),
on two calls in a row, the second with a smaller problem (the scratch keeps the first call's
contents, which the second must not see).  A product whose workspace exceeds the scratch
is rejected with an error when the exact product is requested explicitly, and C is left as
it was."""
import numpy as np
import pytest

import ddgen
import mwblas

pytestmark = pytest.mark.gpu


def _problem(rng, gpu, m, n, k, ta, tb, kind):
    import torch
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = torch.from_numpy(ddgen.real_matrix(rng, ar, ac, ar + 3, kind=kind)).to(gpu)
    B = torch.from_numpy(ddgen.real_matrix(rng, br, bc, br + 1, kind=kind)).to(gpu)
    C0 = torch.from_numpy(ddgen.real_matrix(rng, m, n, m + 2)).to(gpu)
    return dict(m=m, n=n, k=k, ta=ta, tb=tb, A=A, lda=ar + 3, B=B, ldb=br + 1, C0=C0, ldc=m + 2,
                al=ddgen.real_scalar(rng, scale=2), be=ddgen.real_scalar(rng, scale=2))


def _run(p, entry, path):
    import torch
    C = p["C0"].clone()
    entry(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["al"], p["A"], p["lda"], p["B"], p["ldb"],
          p["be"], C, p["ldc"], path=path, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return C


def _existing(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, path, stream):
    mwblas.gemm_device(False, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc, path=path,
                       stream=stream)


@pytest.mark.parametrize("path", [mwblas.PATH_OZAKI, mwblas.PATH_PLAIN, mwblas.PATH_AUTO],
                         ids=["ozaki", "plain", "auto"])
def test_two_calls_match_the_allocating_entry(gpu, path):
    import torch
    rng = np.random.default_rng(1400 + path)
    first = _problem(rng, gpu, 700, 650, 900, "N", "T", "wide")
    second = _problem(rng, gpu, 130, 70, 333, "T", "N", "uniform")
    for name, p in (("first 700x650x900", first), ("second 130x70x333", second)):
        print(f"no-allocation entry, path {path}, {name}", flush=True)
        got = _run(p, mwblas.gemm_device_noalloc, path)
        want = _run(p, _existing, path)
        assert torch.equal(got, want), name


def test_data_dependent_count_in_the_scratch(gpu):
    """With the data-dependent modulus count on for every product, the no-allocation entry
    point chooses the count from the norm bounds it keeps in the scratch memory, with every
    count's constants in the scratch's head: the same bits as the allocating entry point,
    on two calls in a row (the second smaller), and fewer moduli than the worst case for
    the wide-exponent first call."""
    import torch
    rng = np.random.default_rng(1420)
    first = _problem(rng, gpu, 700, 650, 900, "N", "T", "wide")
    second = _problem(rng, gpu, 130, 70, 333, "T", "N", "uniform")
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS)
    try:
        for name, p in (("first 700x650x900", first), ("second 130x70x333", second)):
            print(f"no-allocation entry, data-dependent count, {name}", flush=True)
            got = _run(p, mwblas.gemm_device_noalloc, mwblas.PATH_OZAKI)
            kept = mwblas.ozaki_last_moduli()
            want = _run(p, _existing, mwblas.PATH_OZAKI)
            assert mwblas.ozaki_last_moduli() == kept
            assert torch.equal(got, want), name
            if p is first:
                assert kept < mwblas.ozaki_config(p["k"])[0]
    finally:
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)


def test_workspace_beyond_the_scratch_is_rejected(gpu):
    import torch
    rng = np.random.default_rng(1410)
    p = _problem(rng, gpu, 1536, 1536, 1536, "N", "N", "uniform")
    C = p["C0"].clone()
    print("no-allocation entry, exact product 1536^3 (beyond the scratch): expect an error",
          flush=True)
    with pytest.raises(mwblas.MwgemmError, match="scratch"):
        mwblas.gemm_device_noalloc("N", "N", 1536, 1536, 1536, p["al"], p["A"], p["lda"],
                                   p["B"], p["ldb"], p["be"], C, p["ldc"],
                                   path=mwblas.PATH_OZAKI)
    torch.cuda.synchronize()
    assert torch.equal(C, p["C0"])
    # AUTO computes it on the plain path instead
    got = _run(p, mwblas.gemm_device_noalloc, mwblas.PATH_AUTO)
    want = _run(p, _existing, mwblas.PATH_PLAIN)
    assert torch.equal(got, want)
