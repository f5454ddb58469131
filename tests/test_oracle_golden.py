"""Pin the from-scratch binary128 oracle (oracle/q128_oracle.c) against the reference's own
binary128 GEMMs qgemm_/xgemm_ (src/epblas-parallel/kind16/qgemm_parallel.c,
xgemm_parallel.c) on the committed golden vectors (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "q128_gemm_golden.npz")


def _cases(prefix, count):
    g = np.load(GOLD)
    for i in range(int(g[count])):
        p = f"{prefix}{i}_"
        yield {k[len(p):]: g[k] for k in g.files if k.startswith(p)}


@pytest.mark.parametrize("case", list(_cases("r", "n_real")), ids=lambda c: bytes(c["trans"]).decode())
def test_q128_real_matches_reference_qgemm(case):
    m, n, k, lda, ldb, ldc = (int(v) for v in case["dims"])
    ta, tb = bytes(case["trans"]).decode()
    _, q = oracle.q128_gemm_dd(False, ta, tb, m, n, k, case["alpha"], case["A"], lda,
                               case["B"], ldb, case["beta"], case["C0"], ldc)
    ref = np.ascontiguousarray(case["Cq"].reshape(-1, 16))
    mine = q.reshape(n, m, 16).reshape(-1, 16)
    L = oracle.lib()
    d = L.q128_max_abs_diff(mine.ctypes.data, ref.ctypes.data, m * n)
    scale = L.q128_max_abs(ref.ctypes.data, m * n)
    # Same binary128 arithmetic, different association: agree to a few ulps of 2^-113.
    assert d <= max(scale, 1.0) * k * 2.0 ** -110, (d, scale)


@pytest.mark.parametrize("case", list(_cases("z", "n_cplx")), ids=lambda c: bytes(c["trans"]).decode())
def test_q128_cplx_matches_reference_xgemm(case):
    m, n, k, lda, ldb, ldc = (int(v) for v in case["dims"])
    ta, tb = bytes(case["trans"]).decode()
    _, q = oracle.q128_gemm_dd(True, ta, tb, m, n, k, case["alpha"], case["A"], lda,
                               case["B"], ldb, case["beta"], case["C0"], ldc)
    ref = np.ascontiguousarray(case["Cq"].reshape(-1, 32)).reshape(-1, 16)
    mine = q.reshape(-1, 16)
    L = oracle.lib()
    d = L.q128_max_abs_diff(mine.ctypes.data, ref.ctypes.data, 2 * m * n)
    scale = L.q128_max_abs(ref.ctypes.data, 2 * m * n)
    assert d <= max(scale, 1.0) * k * 2.0 ** -109, (d, scale)


def test_golden_detects_a_planted_mistake():
    """A plausible oracle mistake (conjugating B when op(B) = 'T') must fail the check."""
    case = next(c for c in _cases("z", "n_cplx") if bytes(c["trans"]).decode() == "CT")
    m, n, k, lda, ldb, ldc = (int(v) for v in case["dims"])
    _, q = oracle.q128_gemm_dd(True, "C", "C", m, n, k, case["alpha"], case["A"], lda,
                               case["B"], ldb, case["beta"], case["C0"], ldc)
    ref = np.ascontiguousarray(case["Cq"].reshape(-1, 32)).reshape(-1, 16)
    L = oracle.lib()
    d = L.q128_max_abs_diff(q.reshape(-1, 16).ctypes.data, ref.ctypes.data, 2 * m * n)
    assert d > 1e-3


def test_ref_library_agrees_when_present():
    """Where oracle/_ref exists (built from the reference sources), re-run one case live."""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libqxref.so not built here")
    import ctypes
    case = next(_cases("r", "n_real"))
    m, n, k, lda, ldb, ldc = (int(v) for v in case["dims"])
    ta, tb = bytes(case["trans"]).decode()
    from golden.make_golden import to_q  # noqa: E402
    qa, qb, qc = to_q(case["A"]), to_q(case["B"]), to_q(case["C0"])
    qal, qbe = to_q(case["alpha"].reshape(1, 2)), to_q(case["beta"].reshape(1, 2))
    i32 = ctypes.c_int32
    args = [i32(v) for v in (m, n, k, lda, ldb, ldc)]
    vp = ctypes.c_void_p
    ref.qgemm_(ta.encode(), tb.encode(), *[ctypes.byref(a) for a in args[:3]],
               vp(qal.ctypes.data), vp(qa.ctypes.data), ctypes.byref(args[3]),
               vp(qb.ctypes.data), ctypes.byref(args[4]), vp(qbe.ctypes.data),
               vp(qc.ctypes.data), ctypes.byref(args[5]))
    assert np.array_equal(qc[:, :m, :], case["Cq"])
