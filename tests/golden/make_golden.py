"""Write tests/golden/q128_gemm_golden.npz: input/output vectors of the reference's own
binary128 GEMMs, qgemm_ and xgemm_ (src/epblas-parallel/kind16/qgemm_parallel.c:39-162,
xgemm_parallel.c:42-165), compiled from their sources by ``make -C oracle ref`` into
oracle/_ref/libqxref.so.  Run in a container where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

Each case stores seeded DD inputs (A, B, C0 with padded leading dimensions, alpha, beta)
and the reference's binary128 result C (16 raw bytes per real component).  The fixture
pins oracle/q128_oracle.c (the from-scratch binary128 oracle) against the reference's
implementation on the same inputs.  The reference has no golden vectors of its own for
this path; these are the nearest thing it can produce.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "epblas-parallel_amd"))

import ddgen  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "q128_gemm_golden.npz")


def to_q(dd):
    """DD array (..., 2) -> binary128 raw bytes (..., 16)."""
    flat = np.ascontiguousarray(dd.reshape(-1, 2))
    out = np.zeros(flat.shape[0] * 16, dtype=np.uint8)
    oracle.lib().q128_from_dd(flat.ctypes.data_as(ctypes.c_void_p), flat.shape[0],
                              out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(dd.shape[:-1] + (16,))


def zto_q(z):
    """complex DD (..., 4) -> interleaved binary128 re/im (..., 32)."""
    re = to_q(z[..., 0:2])
    im = to_q(z[..., 2:4])
    return np.concatenate([re, im], axis=-1)


def main():
    oracle.build(ref=True)
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libqxref.so not built (needs /root/reference)")
    rng = np.random.default_rng(20240617)
    i32 = ctypes.c_int32
    cases = {}
    real_shapes = [("N", "N", 7, 5, 9), ("T", "N", 16, 11, 33), ("N", "T", 13, 17, 130),
                   ("T", "T", 9, 8, 21), ("N", "N", 1, 1, 1), ("C", "N", 12, 6, 140),
                   ("N", "N", 24, 20, 256), ("T", "N", 5, 31, 3)]
    for idx, (ta, tb, m, n, k) in enumerate(real_shapes):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = ar + idx % 3, br + (idx * 2) % 5, m + idx % 4
        A = ddgen.real_matrix(rng, ar, ac, lda)
        B = ddgen.real_matrix(rng, br, bc, ldb)
        C0 = ddgen.real_matrix(rng, m, n, ldc)
        alpha = ddgen.real_scalar(rng, scale=2.0)
        beta = ddgen.real_scalar(rng, scale=2.0) if idx % 4 else np.zeros(2)
        qa, qb, qc = to_q(A), to_q(B), to_q(C0)
        qal, qbe = to_q(alpha.reshape(1, 2)), to_q(beta.reshape(1, 2))
        qc_out = qc.copy()
        args = [i32(v) for v in (m, n, k, lda, ldb, ldc)]
        ref.qgemm_(ta.encode(), tb.encode(), ctypes.byref(args[0]), ctypes.byref(args[1]),
                   ctypes.byref(args[2]), qal.ctypes.data_as(ctypes.c_void_p),
                   qa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(args[3]),
                   qb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(args[4]),
                   qbe.ctypes.data_as(ctypes.c_void_p), qc_out.ctypes.data_as(ctypes.c_void_p),
                   ctypes.byref(args[5]))
        p = f"r{idx}_"
        cases[p + "dims"] = np.array([m, n, k, lda, ldb, ldc], dtype=np.int64)
        cases[p + "trans"] = np.frombuffer((ta + tb).encode(), dtype=np.uint8).copy()
        cases[p + "A"], cases[p + "B"], cases[p + "C0"] = A, B, C0
        cases[p + "alpha"], cases[p + "beta"] = alpha, beta
        cases[p + "Cq"] = qc_out[:, :m, :].copy()
    cplx_shapes = [("N", "N", 6, 7, 5), ("C", "N", 11, 9, 17), ("N", "C", 8, 13, 140),
                   ("T", "C", 5, 6, 33), ("C", "T", 10, 4, 9), ("C", "C", 7, 7, 12)]
    for idx, (ta, tb, m, n, k) in enumerate(cplx_shapes):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = ar + idx % 2, br + idx % 3, m + 1
        A = ddgen.cplx_matrix(rng, ar, ac, lda)
        B = ddgen.cplx_matrix(rng, br, bc, ldb)
        C0 = ddgen.cplx_matrix(rng, m, n, ldc)
        alpha = ddgen.cplx_scalar(rng, scale=2.0)
        beta = ddgen.cplx_scalar(rng, scale=2.0) if idx % 3 else np.zeros(4)
        qa, qb, qc = zto_q(A), zto_q(B), zto_q(C0)
        qal, qbe = zto_q(alpha.reshape(1, 4)), zto_q(beta.reshape(1, 4))
        qc_out = qc.copy()
        args = [i32(v) for v in (m, n, k, lda, ldb, ldc)]
        ref.xgemm_(ta.encode(), tb.encode(), ctypes.byref(args[0]), ctypes.byref(args[1]),
                   ctypes.byref(args[2]), qal.ctypes.data_as(ctypes.c_void_p),
                   qa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(args[3]),
                   qb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(args[4]),
                   qbe.ctypes.data_as(ctypes.c_void_p), qc_out.ctypes.data_as(ctypes.c_void_p),
                   ctypes.byref(args[5]))
        p = f"z{idx}_"
        cases[p + "dims"] = np.array([m, n, k, lda, ldb, ldc], dtype=np.int64)
        cases[p + "trans"] = np.frombuffer((ta + tb).encode(), dtype=np.uint8).copy()
        cases[p + "A"], cases[p + "B"], cases[p + "C0"] = A, B, C0
        cases[p + "alpha"], cases[p + "beta"] = alpha, beta
        cases[p + "Cq"] = qc_out[:, :m, :].copy()
    cases["n_real"] = np.array(len(real_shapes))
    cases["n_cplx"] = np.array(len(cplx_shapes))
    np.savez_compressed(OUT, **cases)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
