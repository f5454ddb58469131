"""Exact-product calls captured into a HIP graph keep a workspace of their own.

A captured call bakes the workspace's address into the graph.  A later eager call that
needs a larger workspace must not free or overwrite that buffer: it moves to a new one, and
the graph keeps the old.  The test captures a small exact-product GEMM, makes a larger
eager call (which used to free the captured buffer), then replays the graph several times,
with eager calls of the library on another stream between and beside the replays, and
checks every replay and every eager call against the same call made alone.  With the
data-dependent modulus count on for every product, the graph also holds the row norm
bounds' clearing and atomics, the choice of the count and the GEMMs and reconstruction that
read it on the device."""
import numpy as np
import pytest

import ddgen
import mwblas

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("estimate", ["large", "always"])
def test_replays_survive_a_larger_eager_call(gpu, estimate):
    import torch
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS if estimate == "always"
                              else mwblas.ESTIMATE_LARGE)
    try:
        _replays_survive_a_larger_eager_call(gpu)
    finally:
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)


def _replays_survive_a_larger_eager_call(gpu):
    import torch
    rng = np.random.default_rng(1301)
    small, big = (300, 260, 410), (1100, 900, 1300)

    def problem(m, n, k):
        A = torch.from_numpy(ddgen.real_matrix(rng, m, k)).to(gpu)
        B = torch.from_numpy(ddgen.real_matrix(rng, k, n)).to(gpu)
        return A, B, torch.zeros((n, m, 2), dtype=torch.float64, device=gpu)

    al, be = ddgen.real_scalar(rng, scale=2), np.zeros(2)    # beta = 0: C is overwritten
    (m1, n1, k1), (m2, n2, k2) = small, big
    A1, B1, C1 = problem(m1, n1, k1)
    A2, B2, C2 = problem(m2, n2, k2)

    def call(m, n, k, A, B, C, stream):
        mwblas.gemm_device(False, "N", "N", m, n, k, al, A, m, B, k, be, C, m,
                           path=mwblas.PATH_OZAKI, stream=stream)

    cur = torch.cuda.current_stream()
    call(m1, n1, k1, A1, B1, C1, cur.cuda_stream)      # warm-up: workspace for the small call
    torch.cuda.synchronize()
    ref1 = C1.clone()

    side = torch.cuda.Stream()
    side.wait_stream(cur)
    graph = torch.cuda.CUDAGraph()
    print("capturing an exact-product 300x260x410 call", flush=True)
    with torch.cuda.graph(graph, stream=side):
        call(m1, n1, k1, A1, B1, C1, side.cuda_stream)
    torch.cuda.synchronize()

    print("eager exact-product 1100x900x1300 call (grows the workspace)", flush=True)
    call(m2, n2, k2, A2, B2, C2, cur.cuda_stream)
    torch.cuda.synchronize()
    ref2 = C2.clone()

    other = torch.cuda.Stream()
    for rep in range(4):
        C1.fill_(7.0)
        C2.fill_(7.0)
        torch.cuda.synchronize()
        graph.replay()
        # an eager call on another stream while the replay may still run
        call(m2, n2, k2, A2, B2, C2, other.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(C1, ref1), f"replay {rep}"
        assert torch.equal(C2, ref2), f"eager call beside replay {rep}"
    del graph
