"""Pin the CPU restatement of the reference triangular-output GEMMs mgemmtr_/wgemmtr_
(oracle/ddgemm_port.c, port_mgemmtr / port_wgemmtr).

What pins what:
* structure: off the nb x nb diagonal triangles, every UPLO element goes through the
  serial GEMM with beta = 1 after the beta scaling (mgemmtr_serial.cpp:146-166), i.e. the
  same per-element sequence as the full GEMM, so those elements equal the (already pinned)
  port_mgemm result bit for bit; elements of the diagonal triangles take the scalar path
  (:48-95), which rounds differently, so differences may appear there and only there;
  the complementary triangle and the padding rows are never written;
* accuracy: the reference fuzz tolerance max(10,4K)*2^-104 normwise against the binary128
  oracle over fuzz-drawn cases (fuzz_gemmtr_real_body.fypp: n, k log-uniform in [1,256],
  random uplo and trans, fuzzed alpha/beta, padded leading dimensions);
* exactness: integer operands give the exact integer product on the triangle;
* driver special cases (alpha == 0 or k == 0 scale only the triangle, beta == 0 clears it,
  n <= 0 is a no-op; mgemmtr_parallel.cpp:43-55).
Bitwise parity of the scalar diagonal path with the third-party multifloats operators
themselves is unpinned, as for the GEMM's beta pre-pass.
"""
import numpy as np
import pytest

import ddgen
import oracle

TOL_EPS = 2.0 ** -104


def _tri_masks(n, upper, nb=8):
    """Boolean (cols, rows) masks of the UPLO triangle and of its nb-block diagonal part."""
    j = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    jc = (j // nb) * nb
    if upper:
        tri = i <= j
        diag = (i >= jc) & (i <= j)
    else:
        tri = i >= j
        diag = (i >= j) & (i < np.minimum(jc + nb, n))
    return tri, diag


def _case(rng, cplx, nmax=256, kmax=256):
    n, k = ddgen.rand_int_log(rng, 1, nmax), ddgen.rand_int_log(rng, 1, kmax)
    uplo = "U" if rng.random() < 0.5 else "L"
    letters = "NTC"
    ta, tb = letters[int(rng.integers(0, 3))], letters[int(rng.integers(0, 3))]
    alpha, beta = ddgen.fuzz_alpha_beta(rng, cplx)
    ar, ac = (n, k) if ta == "N" else (k, n)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ddgen.rand_pad(rng), br + ddgen.rand_pad(rng), n + ddgen.rand_pad(rng)
    gen = ddgen.cplx_matrix if cplx else ddgen.real_matrix
    return dict(uplo=uplo, ta=ta, tb=tb, n=n, k=k, alpha=alpha, beta=beta,
                A=gen(rng, ar, ac, lda), lda=lda, B=gen(rng, br, bc, ldb), ldb=ldb,
                C0=gen(rng, n, n, ldc), ldc=ldc)


def _run_tr(c, cplx, C):
    f = oracle.port_wgemmtr if cplx else oracle.port_mgemmtr
    f(c["uplo"], c["ta"], c["tb"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], c["B"],
      c["ldb"], c["beta"], C, c["ldc"])


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemmtr", "wgemmtr"])
def test_rectangle_is_the_gemm_and_differences_stay_in_the_diagonal_blocks(cplx):
    rng = np.random.default_rng(70 + cplx)
    gemm = oracle.port_wgemm if cplx else oracle.port_mgemm
    diag_diffs = 0
    for _ in range(40):
        c = _case(rng, cplx)
        n = c["n"]
        Ct, Cf = c["C0"].copy(), c["C0"].copy()
        _run_tr(c, cplx, Ct)
        gemm(c["ta"], c["tb"], n, n, c["k"], c["alpha"], c["A"], c["lda"], c["B"], c["ldb"],
             c["beta"], Cf, c["ldc"])
        tri, diag = _tri_masks(n, c["uplo"] == "U")
        same = np.all(Ct[:, :n, :] == Cf[:, :n, :], axis=-1) | (
            np.isnan(Ct[:, :n, :]).any(-1) & np.isnan(Cf[:, :n, :]).any(-1))
        # off-triangle: untouched; rectangle: bitwise the GEMM; padding rows: untouched
        assert np.array_equal(Ct[:, :n, :][~tri], c["C0"][:, :n, :][~tri], equal_nan=True)
        assert np.all(same[tri & ~diag]), (c["uplo"], c["ta"], c["tb"], n, c["k"])
        assert np.array_equal(Ct[:, n:, :], c["C0"][:, n:, :], equal_nan=True)
        diag_diffs += int(np.sum(~same & diag))
    # the diagonal triangles really take the other (scalar, faithful) path
    assert diag_diffs > 0


@pytest.mark.parametrize("cplx", [False, True], ids=["mgemmtr", "wgemmtr"])
def test_fuzz_against_binary128(cplx):
    rng = np.random.default_rng(80 + cplx)
    for _ in range(30):
        c = _case(rng, cplx)
        n = c["n"]
        C = c["C0"].copy()
        _run_tr(c, cplx, C)
        tri, _ = _tri_masks(n, c["uplo"] == "U")
        cols, rows = np.nonzero(tri)
        r = oracle.q128_gemm_err(cplx, c["ta"], c["tb"], n, n, c["k"], c["alpha"], c["A"],
                                 c["lda"], c["B"], c["ldb"], c["beta"], c["C0"], c["ldc"], C,
                                 c["ldc"], entries=(rows, cols))
        tol = max(10, 4 * c["k"]) * TOL_EPS
        assert r["err"] <= tol, (c["uplo"], c["ta"], c["tb"], n, c["k"], r)


def test_integer_operands_give_the_exact_integer_product():
    rng = np.random.default_rng(90)
    for uplo, ta, tb, n, k in [("U", "N", "N", 37, 300), ("L", "T", "N", 64, 129),
                               ("U", "T", "T", 9, 17), ("L", "N", "T", 70, 1)]:
        ar, ac = (n, k) if ta == "N" else (k, n)
        br, bc = (k, n) if tb == "N" else (n, k)
        A = ddgen.real_matrix(rng, ar, ac, kind="integer", bound=64)
        B = ddgen.real_matrix(rng, br, bc, kind="integer", bound=64)
        C0 = ddgen.real_matrix(rng, n, n, kind="integer", bound=1000)
        C = C0.copy()
        oracle.port_mgemmtr(uplo, ta, tb, n, k, np.array([2.0, 0.0]), A, ar, B, br,
                            np.array([-3.0, 0.0]), C, n)
        ia = A[..., 0].T.astype(np.int64)
        ib = B[..., 0].T.astype(np.int64)
        opa = ia if ta == "N" else ia.T
        opb = ib if tb == "N" else ib.T
        exact = 2 * (opa @ opb) - 3 * C0[..., 0].T.astype(np.int64)    # (rows, cols)
        tri, _ = _tri_masks(n, uplo == "U")
        got = C[..., 0].T
        assert np.array_equal(got[tri.T], exact[tri.T].astype(np.float64))
        assert np.array_equal(got[~tri.T], C0[..., 0].T[~tri.T])
        assert np.all(C[..., 1][tri] == 0.0)


def test_driver_special_cases():
    rng = np.random.default_rng(91)
    n, k = 21, 13
    A, B = ddgen.real_matrix(rng, n, k), ddgen.real_matrix(rng, k, n)
    C0 = ddgen.real_matrix(rng, n, n, ld=n + 2)
    C0[5, 3, 0] = np.nan          # in the lower triangle (row 3 < column 5 is upper)
    C0[3, 5, 0] = np.nan          # column 3, row 5: lower
    for uplo in "UL":
        tri, _ = _tri_masks(n, uplo == "U")
        # alpha == 0: the triangle is scaled by beta, nothing else changes
        C = C0.copy()
        beta = np.array([2.5, 1e-20])
        oracle.port_mgemmtr(uplo, "N", "N", n, k, np.zeros(2), A, n, B, k, beta, C, n + 2)
        C1 = C0.copy()
        oracle.port_mgemm("N", "N", n, n, 0, np.zeros(2), A, n, B, k, beta, C1, n + 2)
        assert np.array_equal(C[:, :n][tri], C1[:, :n][tri], equal_nan=True)
        assert np.array_equal(C[:, :n][~tri], C0[:, :n][~tri], equal_nan=True)
        # beta == 0 clears the triangle (NaN included), k == 0
        C = C0.copy()
        oracle.port_mgemmtr(uplo, "N", "N", n, 0, np.array([1.0, 0.0]), A, n, B, k,
                            np.zeros(2), C, n + 2)
        assert np.all(C[:, :n][tri] == 0.0)
        assert np.array_equal(C[:, :n][~tri], C0[:, :n][~tri], equal_nan=True)
        # beta == 1 with alpha == 0: untouched; n == 0: untouched
        for nn, al, be in [(n, np.zeros(2), np.array([1.0, 0.0])),
                           (0, np.array([1.0, 0.0]), np.zeros(2))]:
            C = C0.copy()
            oracle.port_mgemmtr(uplo, "N", "N", nn, k, al, A, n, B, k, be, C, n + 2)
            assert np.array_equal(C, C0, equal_nan=True)
