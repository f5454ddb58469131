"""Exact-product (Ozaki scheme II, INT8 MFMA) GPU path.

Bars: the INT8 residue kernel is exact (bitwise against integer arithmetic); every DD GEMM
result is within normwise error 1e-30 of the binary128 oracle (and within the analytic
bound 4*2^-b of the configured width b); integer operands give the exact integer product."""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

BAR = 1e-30


@pytest.fixture
def ozaki_path():
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    yield
    mwblas.set_path(prev)


def _assert_symmetric_residues(got, expect, m):
    """The residue GEMMs store v - m*rint(v/m): congruent to the exact product modulo m, in
    [-(m-1)/2, (m-1)/2] (odd m) or [-128, 128] (m = 256, whose ties give +-128 = byte 0x80,
    read back as -128)."""
    assert np.array_equal(got % m, expect)
    h = 128 if m == 256 else (m - 1) // 2
    assert got.min() >= -h and got.max() <= h


@pytest.mark.parametrize("mod_index,variant,kp", [
    (0, 1, 320), (1, 1, 320), (5, 1, 320), (28, 1, 320), (39, 1, 320),
    (0, 2, 320), (28, 3, 192),
    # the DMA-ring kernel: fewer, as many and more 64-deep stages than its ring holds (4);
    # 6, 7, 10 and 17 stages hand over from the stage pairs to the single-buffer stages
    # after an even or an odd number of refills
    (0, 0, 64), (5, 0, 128), (28, 0, 192), (39, 0, 256), (1, 0, 320), (3, 0, 384),
    (7, 0, 448), (11, 0, 640), (0, 0, 1088)])
def test_residue_gemm_kernel_is_exact(gpu, mod_index, variant, kp):
    import torch
    rng = np.random.default_rng(100 + mod_index + 7 * variant + kp)
    mp, np_ = 256, 512
    m_all, _ = mwblas.ozaki_moduli(40)
    m = int(m_all[mod_index])
    lo, hi = (-128, 127) if m == 256 else (-(m - 1) // 2, (m - 1) // 2)
    A = rng.integers(lo, hi + 1, size=(mp, kp), dtype=np.int64)
    B = rng.integers(lo, hi + 1, size=(np_, kp), dtype=np.int64)
    dA = torch.from_numpy(A.astype(np.int8)).to(gpu)
    dB = torch.from_numpy(B.astype(np.int8)).to(gpu)
    dU = torch.zeros((np_, mp), dtype=torch.uint8, device=gpu)
    print(f"residue gemm variant {variant} mod {mod_index} kp {kp}", flush=True)
    mwblas.testing_residue_gemm(dA, dB, dU, mp, np_, kp, mod_index, variant)
    torch.cuda.synchronize()
    expect = (A @ B.T) % m                        # (mp, np); y lives in the CRT limbs
    got = mwblas.residue_products(dU.cpu().numpy(), mp, np_).astype(np.int64).T
    _assert_symmetric_residues(got, expect, m)


def _check(cplx, ta, tb, m, n, k, kind, rng, ent=None):
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + 1, br + 2, m + 3
    A, B = gen(rng, ar, ac, lda, kind=kind), gen(rng, br, bc, ldb, kind=kind)
    C0 = gen(rng, m, n, ldc)
    al = ddgen.cplx_scalar(rng, scale=2) if cplx else ddgen.real_scalar(rng, scale=2)
    be = ddgen.cplx_scalar(rng, scale=2) if cplx else ddgen.real_scalar(rng, scale=2)
    C = C0.copy()
    (mwblas.wgemm if cplx else mwblas.mgemm)(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc)
    assert np.array_equal(C[:, m:, :], C0[:, m:, :])       # padding rows untouched
    r = oracle.q128_gemm_err(cplx, ta, tb, m, n, k, al, A, lda, B, ldb, be, C0, ldc, C, ldc,
                             entries=ent)
    _, bits = mwblas.ozaki_config(k, cplx)
    print(f"{'z' if cplx else 'r'} {ta}{tb} {m}x{n}x{k} {kind}: err={r['err']:.3e}", flush=True)
    assert r["err"] <= BAR
    assert r["err"] <= 4 * 2.0 ** -bits
    return r


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("kind", ["uniform", "wide"])
def test_real_vs_binary128(gpu, ozaki_path, ta, tb, kind):
    rng = np.random.default_rng(11)
    _check(False, ta, tb, 300, 211, 517, kind, rng)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("N", "C"), ("T", "C"), ("C", "T"),
                                   ("C", "C")])
def test_cplx_vs_binary128(gpu, ozaki_path, ta, tb):
    rng = np.random.default_rng(12)
    _check(True, ta, tb, 150, 133, 259, "uniform", rng)


def test_small_and_ragged_shapes(gpu, ozaki_path):
    rng = np.random.default_rng(13)
    for (m, n, k) in [(1, 1, 1), (3, 257, 2), (255, 1, 65), (17, 33, 1000)]:
        _check(False, "N", "N", m, n, k, "wide", rng)
        _check(True, "C", "N", m, n, k, "uniform", rng)


def test_integer_operands(gpu, ozaki_path):
    """Integer operands: the high limbs carry the exact integer product, and the whole
    result is within the normwise bar of it (the reconstruction rounds C' = A'B' through a
    double-double multiply, so low limbs of ~2^-104 relative may remain)."""
    rng = np.random.default_rng(14)
    m, n, k = 70, 90, 600
    A = ddgen.real_matrix(rng, m, k, kind="integer", bound=1000)
    B = ddgen.real_matrix(rng, k, n, kind="integer", bound=1000)
    C = np.zeros((n, m, 2))
    mwblas.mgemm("N", "N", m, n, k, [1, 0], A, m, B, k, [0, 0], C, m)
    exact = A[..., 0].T.astype(np.int64) @ B[..., 0].T.astype(np.int64)
    assert np.array_equal(C[..., 0].T, exact.astype(np.float64))
    opn = k * 1000.0 * 1000.0
    assert np.max(np.abs(C[..., 1])) <= 1e-30 * opn


def test_special_values(gpu, ozaki_path):
    rng = np.random.default_rng(15)
    m, n, k = 40, 30, 50
    A, B = ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n)
    A[:, 5, :] = 0.0                    # a zero row of A
    A[7, 9, 0] = np.nan                 # row 9 of A carries a NaN
    C0 = ddgen.real_matrix(rng, m, n)
    C0[2, 3, 0] = np.nan
    C = C0.copy()
    mwblas.mgemm("N", "N", m, n, k, [1, 0], A, m, B, k, [0, 0], C, m)
    assert np.all(np.isnan(C[:, 9, 0]))                      # NaN row propagates
    assert np.all(C[:, 5, :] == 0.0)                          # zero row is exact zero
    mask = np.ones(m, bool)
    mask[9] = False
    assert np.all(np.isfinite(C[:, mask, :]))                 # beta = 0 cleared C0's NaN


def test_calls_on_two_streams_share_the_workspace_safely(gpu):
    """The exact-product workspace (residues and residue products) is one per device.  Calls
    alternating between two streams with no host synchronisation in between must each give
    the result of the same call made alone: a call on another stream waits, on the device,
    for the previous call's kernels before overwriting the workspace."""
    import torch
    rng = np.random.default_rng(77)
    m = n = k = 1024
    cases = [(ddgen.real_matrix(rng, m, k), ddgen.real_matrix(rng, k, n),
              ddgen.real_matrix(rng, m, n), ddgen.real_scalar(rng, scale=2),
              ddgen.real_scalar(rng, scale=2)) for _ in range(4)]
    dev = [tuple(torch.from_numpy(x).to(gpu) for x in c[:3]) for c in cases]
    ref = []
    for (dA, dB, dC0), (_, _, _, al, be) in zip(dev, cases):
        dC = dC0.clone()
        mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, m, dB, k, be, dC, m,
                           path=mwblas.PATH_OZAKI, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        ref.append(dC.cpu().numpy())
    streams = (torch.cuda.Stream(), torch.cuda.Stream())   # default priority
    outs = [torch.empty_like(d[2]) for d in dev]
    for rep in range(3):
        print(f"two-stream round {rep}", flush=True)
        for i, ((dA, dB, dC0), (_, _, _, al, be)) in enumerate(zip(dev, cases)):
            st = streams[i % 2]
            with torch.cuda.stream(st):
                outs[i].copy_(dC0)
                mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, m, dB, k, be, outs[i], m,
                                   path=mwblas.PATH_OZAKI, stream=st.cuda_stream)
        torch.cuda.synchronize()
        for i in range(len(cases)):
            assert np.array_equal(outs[i].cpu().numpy(), ref[i]), (rep, i)


def test_8192_wide_exponents_vs_binary128_and_plain(gpu):
    """The headline workload: M=N=K=8192, hi = +-U[1,2)*2^U_int[-20,20], alpha=1, beta=0.
    Ozaki and plain GPU results: 256 sampled entries each within 1e-30 of binary128."""
    import torch
    rng = np.random.default_rng(8192)
    m = n = k = 8192
    A = ddgen.real_matrix(rng, m, k, kind="wide")
    B = ddgen.real_matrix(rng, k, n, kind="wide")
    C0 = np.zeros((n, m, 2))
    al, be = np.array([1.0, 0.0]), np.zeros(2)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    ent = (rng.integers(0, m, 256), rng.integers(0, n, 256))
    for path in (mwblas.PATH_OZAKI, mwblas.PATH_PLAIN):
        dC = torch.zeros((n, m, 2), dtype=torch.float64, device=gpu)
        mwblas.gemm_device(False, "N", "N", m, n, k, al, dA, m, dB, k, be, dC, m, path=path)
        torch.cuda.synchronize()
        C = dC.cpu().numpy()
        r = oracle.q128_gemm_err(False, "N", "N", m, n, k, al, A, m, B, k, be, C0, m, C, m,
                                 entries=ent)
        print(f"8192 path={path} err={r['err']:.3e}", flush=True)
        assert r["err"] <= BAR


@pytest.mark.parametrize("m,n,k,ta,tb", [(4096, 4096, 1000, "N", "N"), (4352, 4096, 256, "T", "N"),
                                         (4096, 4608, 2048, "N", "T")])
def test_ring_grouped_moduli_vs_binary128(gpu, m, n, k, ta, tb):
    """Real products of depth <= 2048 on the 256 x 256 ring tiles run groups of 8 moduli per
    workgroup as one ring sequence (29 = 8 + 8 + 8 + 5).  Shapes large enough for the ring
    tiles, ragged rows, both transposes (at depth 256 the row-contiguous op(A) of 'T' and
    op(B) of 'N' take maxima and split in one kernel); 192 sampled entries against
    binary128."""
    import torch
    rng = np.random.default_rng(m + n + k)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A, B = ddgen.real_matrix(rng, ar, ac, kind="wide"), ddgen.real_matrix(rng, br, bc, kind="wide")
    C0 = ddgen.real_matrix(rng, m, n)
    al, be = ddgen.real_scalar(rng, scale=2), ddgen.real_scalar(rng, scale=2)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    dC = torch.from_numpy(C0).to(gpu)
    print(f"ring grouped moduli {m}x{n}x{k} {ta}{tb}", flush=True)
    mwblas.gemm_device(False, ta, tb, m, n, k, al, dA, ar, dB, br, be, dC, m,
                       path=mwblas.PATH_OZAKI)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    ent = (rng.integers(0, m, 192), rng.integers(0, n, 192))
    r = oracle.q128_gemm_err(False, ta, tb, m, n, k, al, A, ar, B, br, be, C0, m, C, m,
                             entries=ent)
    print(f"  err={r['err']:.3e}", flush=True)
    assert r["err"] <= BAR


def test_4096_complex_conj_vs_binary128(gpu):
    """Complex workload: M=N=K=4096, op(A) = conjugate transpose, op(B) = N, random complex
    DD alpha/beta; plain and Ozaki, 128 sampled entries against binary128."""
    import torch
    rng = np.random.default_rng(4097)
    m = n = k = 4096
    A = ddgen.cplx_matrix(rng, k, m)
    B = ddgen.cplx_matrix(rng, k, n)
    C0 = ddgen.cplx_matrix(rng, m, n)
    al, be = ddgen.cplx_scalar(rng), ddgen.cplx_scalar(rng)
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    ent = (rng.integers(0, m, 128), rng.integers(0, n, 128))
    for path in (mwblas.PATH_OZAKI, mwblas.PATH_PLAIN):
        dC = torch.from_numpy(C0).to(gpu)
        mwblas.gemm_device(True, "C", "N", m, n, k, al, dA, k, dB, k, be, dC, m, path=path)
        torch.cuda.synchronize()
        C = dC.cpu().numpy()
        r = oracle.q128_gemm_err(True, "C", "N", m, n, k, al, A, k, B, k, be, C0, m, C, m,
                                 entries=ent)
        print(f"4096 complex path={path} err={r['err']:.3e}", flush=True)
        assert r["err"] <= BAR


@pytest.mark.parametrize("variant,kp", [(0, 70016), (1, 70016), (2, 40000), (3, 40000)])
def test_residue_gemm_deeper_than_one_accumulation(gpu, variant, kp):
    """Depth beyond 32768: every residue-GEMM tile reduces its INT32 sums modulo m on the
    way, so the stored residue is still exact (integer reference in float64, exact below
    2^53)."""
    import torch
    rng = np.random.default_rng(140 + variant)
    mp, np_ = 256, 512
    m_all, _ = mwblas.ozaki_moduli(40)
    mod_index = (1, 5, 28, 39)[variant]
    m = int(m_all[mod_index])
    h = (m - 1) // 2
    A = rng.integers(-h, h + 1, size=(mp, kp), dtype=np.int64)
    B = rng.integers(-h, h + 1, size=(np_, kp), dtype=np.int64)
    dA = torch.from_numpy(A.astype(np.int8)).to(gpu)
    dB = torch.from_numpy(B.astype(np.int8)).to(gpu)
    dU = torch.zeros((np_, mp), dtype=torch.uint8, device=gpu)
    print(f"deep residue gemm variant {variant} mod {m} kp {kp}", flush=True)
    mwblas.testing_residue_gemm(dA, dB, dU, mp, np_, kp, mod_index, variant)
    torch.cuda.synchronize()
    expect = (A.astype(np.float64) @ B.T.astype(np.float64)).astype(np.int64) % m
    got = mwblas.residue_products(dU.cpu().numpy(), mp, np_).astype(np.int64).T
    _assert_symmetric_residues(got, expect, m)


@pytest.mark.parametrize("variant", [0, 1])
def test_residue_gemm_sum_beyond_int32(gpu, variant):
    """127 * 127 over a depth of 135168 sums to 2.18e9 > 2^31: an INT32 accumulator without
    the modular reduction would wrap (and 2^32 = 1 mod 255 would change the residue)."""
    import torch
    mp, np_, kp, mod_index = 256, 256, 135168, 1
    m = int(mwblas.ozaki_moduli(40)[0][mod_index])
    assert m == 255
    dA = torch.full((mp, kp), 127, dtype=torch.int8, device=gpu)
    dB = torch.full((np_, kp), 127, dtype=torch.int8, device=gpu)
    dB[1::2, :] = -127
    dU = torch.zeros((np_, mp), dtype=torch.uint8, device=gpu)
    print(f"residue gemm variant {variant}, constant +-127 operands, kp {kp}", flush=True)
    mwblas.testing_residue_gemm(dA, dB, dU, mp, np_, kp, mod_index, variant)
    torch.cuda.synchronize()
    got = mwblas.residue_products(dU.cpu().numpy(), mp, np_).astype(np.int64)
    _assert_symmetric_residues(got[0::2], np.full_like(got[0::2], (127 * 127 * kp) % m), m)
    _assert_symmetric_residues(got[1::2], np.full_like(got[1::2], (-127 * 127 * kp) % m), m)


@pytest.mark.parametrize("cplx,k", [(False, 70000), (True, 40000)])
def test_exact_product_deeper_than_65536(gpu, cplx, k):
    """Real k = 70000 and complex k = 40000 (2k = 80000) stay on the exact-product path and
    within 1e-30 of binary128 on sampled entries."""
    import torch
    rng = np.random.default_rng(150 + cplx)
    m, n = (256, 192) if not cplx else (160, 128)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    scal = ddgen.cplx_scalar if cplx else ddgen.real_scalar
    A, B, C0 = gen(rng, k, m), gen(rng, k, n), gen(rng, m, n)
    al, be = scal(rng, scale=2), scal(rng, scale=2)
    ta = "C" if cplx else "T"
    dA, dB, dC = (torch.from_numpy(x).to(gpu) for x in (A, B, C0))
    mwblas.timer_reset()
    mwblas.timer_enable(True)
    try:
        mwblas.gemm_device(cplx, ta, "N", m, n, k, al, dA, k, dB, k, be, dC, m,
                           path=mwblas.PATH_OZAKI)
        torch.cuda.synchronize()
    finally:
        mwblas.timer_enable(False)
    kt = mwblas.timer_read()
    assert kt.get("oz_gemm", (0, 0))[1] == 1 and kt.get("plain_real", (0, 0))[1] == 0
    assert kt.get("plain_cplx", (0, 0))[1] == 0
    ent = (rng.integers(0, m, 96), rng.integers(0, n, 96))
    r = oracle.q128_gemm_err(cplx, ta, "N", m, n, k, al, A, k, B, k, be, C0, m,
                             dC.cpu().numpy(), m, entries=ent)
    nmod, bits = mwblas.ozaki_config(k, cplx)
    print(f"deep {'complex' if cplx else 'real'} k={k}: {nmod} moduli, err={r['err']:.3e}",
          flush=True)
    assert r["err"] <= BAR
    assert r["err"] <= 4 * 2.0 ** -bits


ONE, ZERO = np.array([1.0, 0.0]), np.zeros(2)


@pytest.fixture
def estimate_always():
    """The data-dependent count on every real product (by default only on products of at
    least 2^33 multiply-adds); the count-offset test hook reset afterwards."""
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS)
    yield
    mwblas.ozaki_set_estimate(mwblas.ESTIMATE_LARGE)
    mwblas.testing_ozaki_count_offset(0)


def _real_nn(gpu, A, B, m, n, k):
    import torch
    dA, dB = torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu)
    dC = torch.zeros((n, m, 2), dtype=torch.float64, device=gpu)
    mwblas.gemm_device(False, "N", "N", m, n, k, [1.0, 0.0], dA, m, dB, k, [0.0, 0.0], dC, m,
                       path=mwblas.PATH_OZAKI)
    torch.cuda.synchronize()
    return dC.cpu().numpy()


def test_estimated_count_saves_a_modulus_and_one_fewer_fails(gpu, estimate_always):
    """Data-dependent modulus count on the 512^3 workload (alpha = 1, beta = 0, uniform
    entries): the Cauchy-Schwarz bound of the scaled rows keeps one modulus fewer than the
    worst case k 2^(2b); every entry is within 1e-30 of binary128; a second call is bitwise
    equal (the bounds are integer sums); and one more modulus removed (test hook) wraps the
    reconstruction, so the kept count is not more than needed by a wide margin."""
    rng = np.random.default_rng(512)
    m = n = k = 512
    A = ddgen.real_matrix(rng, m, k)
    B = ddgen.real_matrix(rng, k, n)
    C0 = np.zeros((n, m, 2))
    worst, bits = mwblas.ozaki_config(k)
    assert bits == 102
    C = _real_nn(gpu, A, B, m, n, k)
    kept = mwblas.ozaki_last_moduli()
    r = oracle.q128_gemm_err(False, "N", "N", m, n, k, ONE, A, m, B, k, ZERO, C0,
                             m, C, m)
    print(f"512^3 uniform: kept {kept} of {worst} moduli, err={r['err']:.3e}", flush=True)
    assert kept == worst - 1
    assert r["err"] <= BAR and r["err"] <= 4 * 2.0 ** -bits
    assert np.array_equal(_real_nn(gpu, A, B, m, n, k), C)
    mwblas.testing_ozaki_count_offset(-1)
    Cbad = _real_nn(gpu, A, B, m, n, k)
    assert mwblas.ozaki_last_moduli() == kept - 1
    rb = oracle.q128_gemm_err(False, "N", "N", m, n, k, ONE, A, m, B, k, ZERO, C0,
                              m, Cbad, m)
    print(f"  with {kept - 1} moduli: err={rb['err']:.3e}", flush=True)
    assert rb["err"] > 1e-3


def test_estimated_count_keeps_the_worst_case_when_it_is_attained(gpu, estimate_always):
    """Adversarial operands: every entry 1 (each row's maximum and every scaled integer
    equal 2^(b-1)), so C'_ij = k 2^(2b-2) and the Cauchy-Schwarz bound is attained.  At
    k = 2048 the bound needs the worst-case count (29); the exact result 2048 comes out, and
    with one modulus fewer (test hook) C' exceeds M/2 and the result is wrong."""
    m, n, k = 256, 320, 2048
    A = np.zeros((k, m, 2))
    A[..., 0] = 1.0
    B = np.zeros((n, k, 2))
    B[..., 0] = 1.0
    worst, bits = mwblas.ozaki_config(k)
    C = _real_nn(gpu, A, B, m, n, k)
    kept = mwblas.ozaki_last_moduli()
    print(f"all-ones k={k}: kept {kept} of {worst} moduli", flush=True)
    assert kept == worst
    assert np.all(C[..., 0] == float(k)) and np.all(C[..., 1] == 0.0)
    mwblas.testing_ozaki_count_offset(-1)
    Cbad = _real_nn(gpu, A, B, m, n, k)
    assert mwblas.ozaki_last_moduli() == worst - 1
    assert not np.any(np.abs(Cbad[..., 0] - k) < 1.0)


def test_estimated_and_worst_case_counts_agree_with_binary128(gpu, estimate_always):
    """Wide-exponent operands, ragged shape: the data-dependent count keeps fewer moduli than
    the worst case, and both settings are within 1e-30 of binary128."""
    rng = np.random.default_rng(4242)
    m, n, k = 700, 520, 3000
    A = ddgen.real_matrix(rng, m, k, kind="wide")
    B = ddgen.real_matrix(rng, k, n, kind="wide")
    C0 = np.zeros((n, m, 2))
    worst, _ = mwblas.ozaki_config(k)
    errs = {}
    for est in (True, False):
        mwblas.ozaki_set_estimate(mwblas.ESTIMATE_ALWAYS if est else mwblas.ESTIMATE_OFF)
        C = _real_nn(gpu, A, B, m, n, k)
        if est:
            kept = mwblas.ozaki_last_moduli()
        r = oracle.q128_gemm_err(False, "N", "N", m, n, k, ONE, A, m, B, k, ZERO, C0, m, C, m,
                                 entries=(rng.integers(0, m, 256), rng.integers(0, n, 256)))
        errs[est] = r["err"]
    print(f"wide {m}x{n}x{k}: kept {kept} of {worst}; err est={errs[True]:.3e} "
          f"worst={errs[False]:.3e}", flush=True)
    assert kept < worst
    assert errs[True] <= BAR and errs[False] <= BAR
