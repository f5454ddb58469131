"""The by-value cores mgemm_serial / wgemm_serial driven the way the reference's blocked
level-3 routines drive them: sequences of trailing updates on interior blocks of larger
host matrices, with the parent leading dimensions, beta = 1, alpha or -1, and block size
nb = 64 (msyrk_serial.cpp:303-305).

Each sequence is replayed twice on copies of the same host arrays: once through the
library (host pointers, staged to the GPU per call) and once through the CPU restatement
of the reference (oracle port_mgemm / port_wgemm, whose per-element operation order is
the one mgemm_serial / wgemm_serial share with mgemm_ / wgemm_).  The whole parent arrays
are compared bit for bit after the sequence, so entries outside each call's block must be
left as they were.  The diagonal-block kernels the reference runs between the updates are
not part of this library and are left out: each call reads what earlier calls wrote, so
the in-place, aliasing and ordering semantics are still exercised.

Call patterns:
* msyrk lower 'N' / upper 'T' (msyrk_serial.cpp:331-356): ('N','T') / ('T','N') blocks of C;
* msyr2k (msyr2k_serial.cpp:346-374): two dependent calls on the same C block;
* mtrsm LLN / LLT (mtrsm_serial.cpp:766-799): alpha = -1 and the B operand and C are
  disjoint row ranges of one array (B_chunk and &B_chunk[ic]);
* msymm right side (msymm_serial.cpp:435-455): ('N','T') and ('N','N') on column blocks;
* whemm left (whemm_serial.cpp:468-482): ('C','N');
* wher2k (wher2k_serial.cpp:489-490): ('N','C') with alpha and conj(alpha).
"""
import numpy as np
import pytest

import ddgen
import mwblas
import oracle

pytestmark = pytest.mark.gpu

NB = 64


def at(X, i, j):
    """Interior pointer to element (i, j) of a column-major DD array (cols, ld, limbs)."""
    return X.reshape(-1, X.shape[2])[j * X.shape[1] + i:]


def _one(cplx):
    return np.array([1.0, 0, 0, 0]) if cplx else np.array([1.0, 0])


def _minus_one(cplx):
    return np.array([-1.0, 0, 0, 0]) if cplx else np.array([-1.0, 0])


def _both(cplx, calls, arrays):
    """Run the call list on two copies of `arrays`: the library and the restatement."""
    lib_arr = {k: v.copy() for k, v in arrays.items()}
    ref_arr = {k: v.copy() for k, v in arrays.items()}
    entry = mwblas.wgemm_serial if cplx else mwblas.mgemm_serial
    port = oracle.port_wgemm if cplx else oracle.port_mgemm
    for n_call, (ta, tb, m, n, k, alpha, a, lda, b, ldb, c, ldc) in enumerate(calls):
        for arr, f in ((lib_arr, entry), (ref_arr, port)):
            f(ta, tb, m, n, k, alpha, at(arr[a[0]], *a[1:]), lda, at(arr[b[0]], *b[1:]), ldb,
              _one(cplx), at(arr[c[0]], *c[1:]), ldc)
    for key in arrays:
        assert np.array_equal(lib_arr[key], ref_arr[key], equal_nan=True), key
    return lib_arr


def _blocks(n):
    return [(jc, min(NB, n - jc)) for jc in range(0, n, NB)]


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T"), ("L", "T"), ("U", "N")])
def test_msyrk_trailing_updates(gpu, uplo, trans):
    rng = np.random.default_rng(800 + "LU".index(uplo) * 2 + "NT".index(trans))
    n, k = 300, 140
    ar, ac = (n, k) if trans == "N" else (k, n)
    lda, ldc = ar + 3, n + 5
    A = ddgen.real_matrix(rng, ar, ac, lda)
    C = ddgen.real_matrix(rng, n, n, ldc)
    alpha = ddgen.real_scalar(rng, scale=2)
    calls = []
    for jc, jb in _blocks(n):
        if uplo == "L":
            j0, rows = jc + jb, n - jc - jb
            if rows <= 0:
                continue
            if trans == "N":
                calls.append(("N", "T", rows, jb, k, alpha, ("A", j0, 0), lda, ("A", jc, 0), lda,
                              ("C", j0, jc), ldc))
            else:
                calls.append(("T", "N", rows, jb, k, alpha, ("A", 0, j0), lda, ("A", 0, jc), lda,
                              ("C", j0, jc), ldc))
        elif jc > 0:
            if trans == "N":
                calls.append(("N", "T", jc, jb, k, alpha, ("A", 0, 0), lda, ("A", jc, 0), lda,
                              ("C", 0, jc), ldc))
            else:
                calls.append(("T", "N", jc, jb, k, alpha, ("A", 0, 0), lda, ("A", 0, jc), lda,
                              ("C", 0, jc), ldc))
    _both(False, calls, {"A": A, "C": C})


def test_msyr2k_two_dependent_updates_per_block(gpu):
    rng = np.random.default_rng(810)
    n, k = 260, 97
    lda, ldb, ldc = n + 1, n + 2, n + 7
    arrays = {"A": ddgen.real_matrix(rng, n, k, lda), "B": ddgen.real_matrix(rng, n, k, ldb),
              "C": ddgen.real_matrix(rng, n, n, ldc)}
    alpha = ddgen.real_scalar(rng, scale=2)
    calls = []
    for jc, jb in _blocks(n):
        j0, rows = jc + jb, n - jc - jb
        if rows > 0:
            calls.append(("N", "T", rows, jb, k, alpha, ("A", j0, 0), lda, ("B", jc, 0), ldb,
                          ("C", j0, jc), ldc))
            calls.append(("N", "T", rows, jb, k, alpha, ("B", j0, 0), ldb, ("A", jc, 0), lda,
                          ("C", j0, jc), ldc))
    _both(False, calls, arrays)


@pytest.mark.parametrize("variant", ["LLN", "LLT"])
def test_mtrsm_updates_alias_rows_of_one_array(gpu, variant):
    rng = np.random.default_rng(820 + (variant == "LLT"))
    m, my_n = 333, 75
    lda, ldb = m + 4, m + 9
    A = ddgen.real_matrix(rng, m, m, lda)
    Bc = ddgen.real_matrix(rng, m, my_n, ldb)
    m_one = _minus_one(False)
    calls = []
    if variant == "LLN":                     # ascending ic, reads rows [0, ic)
        for ic, ib in _blocks(m):
            if ic > 0:
                calls.append(("N", "N", ib, my_n, ic, m_one, ("A", ic, 0), lda, ("B", 0, 0), ldb,
                              ("B", ic, 0), ldb))
    else:                                    # descending ic, reads rows [ic + ib, m)
        for ic, ib in reversed(_blocks(m)):
            rows = m - (ic + ib)
            if rows > 0:
                i0 = ic + ib
                calls.append(("T", "N", ib, my_n, rows, m_one, ("A", i0, ic), lda,
                              ("B", i0, 0), ldb, ("B", ic, 0), ldb))
    out = _both(False, calls, {"A": A, "B": Bc})
    # the first block row is never written
    assert np.array_equal(out["B"][:, :NB] if variant == "LLN" else out["B"][:, m - m % NB:m],
                          (Bc[:, :NB] if variant == "LLN" else Bc[:, m - m % NB:m]),
                          equal_nan=True)


def test_msymm_right_side_column_blocks(gpu):
    rng = np.random.default_rng(830)
    m, n = 190, 230
    lda, ldb, ldc = n + 2, m + 3, m + 1
    arrays = {"A": ddgen.real_matrix(rng, n, n, lda), "B": ddgen.real_matrix(rng, m, n, ldb),
              "C": ddgen.real_matrix(rng, m, n, ldc)}
    alpha = ddgen.real_scalar(rng, scale=2)
    calls = []
    for jc, jb in _blocks(n):
        rows = n - jc - jb
        if jc > 0:
            calls.append(("N", "T", m, jb, jc, alpha, ("B", 0, 0), ldb, ("A", jc, 0), lda,
                          ("C", 0, jc), ldc))
        if rows > 0:
            calls.append(("N", "N", m, jb, rows, alpha, ("B", 0, jc + jb), ldb,
                          ("A", jc + jb, jc), lda, ("C", 0, jc), ldc))
    _both(False, calls, arrays)


def test_whemm_left_conjugate_updates(gpu):
    rng = np.random.default_rng(840)
    m, n = 210, 90
    lda, ldb, ldc = m + 2, m + 1, m + 4
    arrays = {"A": ddgen.cplx_matrix(rng, m, m, lda), "B": ddgen.cplx_matrix(rng, m, n, ldb),
              "C": ddgen.cplx_matrix(rng, m, n, ldc)}
    alpha = ddgen.cplx_scalar(rng, scale=2)
    calls = []
    for ic, ib in _blocks(m):
        rows = m - ic - ib
        if ic > 0:
            calls.append(("N", "N", ib, n, ic, alpha, ("A", ic, 0), lda, ("B", 0, 0), ldb,
                          ("C", ic, 0), ldc))
        if rows > 0:
            calls.append(("C", "N", ib, n, rows, alpha, ("A", ic + ib, ic), lda,
                          ("B", ic + ib, 0), ldb, ("C", ic, 0), ldc))
    _both(True, calls, arrays)


def test_wher2k_alpha_and_conjugate_alpha(gpu):
    rng = np.random.default_rng(850)
    n, k = 200, 70
    lda, ldb, ldc = n + 3, n + 1, n + 2
    arrays = {"A": ddgen.cplx_matrix(rng, n, k, lda), "B": ddgen.cplx_matrix(rng, n, k, ldb),
              "C": ddgen.cplx_matrix(rng, n, n, ldc)}
    alpha = ddgen.cplx_scalar(rng, scale=2)
    alpha_conj = alpha * np.array([1.0, 1.0, -1.0, -1.0])
    calls = []
    for jc, jb in _blocks(n):
        j0, rows = jc + jb, n - jc - jb
        if rows > 0:
            calls.append(("N", "C", rows, jb, k, alpha, ("A", j0, 0), lda, ("B", jc, 0), ldb,
                          ("C", j0, jc), ldc))
            calls.append(("N", "C", rows, jb, k, alpha_conj, ("B", j0, 0), ldb,
                          ("A", jc, 0), lda, ("C", j0, jc), ldc))
    _both(True, calls, arrays)
