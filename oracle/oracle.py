"""ctypes loader for the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module, and only as the checker / the timed CPU reference; the product never does.

Array convention shared with the tests: a column-major DD matrix with leading dimension
``ld`` and ``cols`` columns is a C-contiguous float64 numpy array of shape ``(cols, ld, 2)``
holding ``(hi, lo)``; a complex DD matrix is ``(cols, ld, 4)`` holding
``(re.hi, re.lo, im.hi, im.lo)``.  A DD scalar is a length-2 float64 array, a complex DD
scalar a length-4 one.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_c_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and, where /root/reference exists, the reference binary128
    GEMMs into _ref/)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "-B", "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        _LIB = ctypes.CDLL(path)
        c = ctypes.c_char
        for name in ("port_mgemm", "port_wgemm"):
            f = getattr(_LIB, name)
            f.restype = ctypes.c_int
            f.argtypes = [c, c, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp,
                          _vp, _c_i64, ctypes.c_int]
        for name in ("port_mgemmtr", "port_wgemmtr"):
            f = getattr(_LIB, name)
            f.restype = ctypes.c_int
            f.argtypes = [c, c, c, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp,
                          _c_i64, _c_i64, ctypes.c_int]
        _LIB.port_mgemv.restype = ctypes.c_int
        _LIB.port_mgemv.argtypes = [c, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp,
                                    _c_i64, _c_i64]
        _LIB.port_wgemv.restype = ctypes.c_int
        _LIB.port_wgemv.argtypes = [c, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _vp,
                                    _c_i64]
        for name in ("q128_mgemm_entries", "q128_wgemm_entries"):
            f = getattr(_LIB, name)
            f.restype = None
            f.argtypes = [c, c, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp,
                          _vp, _c_i64, _c_i64, _vp, _vp, _vp, _vp]
        for name in ("q128_mgemm_err", "q128_wgemm_err"):
            f = getattr(_LIB, name)
            f.restype = None
            f.argtypes = [c, c, _c_i64, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp,
                          _vp, _c_i64, _vp, _c_i64, _c_i64, _vp, _vp, _vp]
        for name in ("port_dd_mul_fast", "port_dd_add_fast", "port_dd_mul_exact",
                     "port_dd_add_exact"):
            getattr(_LIB, name).argtypes = [_vp, _vp, _vp]
        for name in ("port_eft_prod", "port_eft_sum"):
            getattr(_LIB, name).argtypes = [ctypes.c_double, ctypes.c_double, _vp, _vp]
        for name in ("q128_check_prod", "q128_check_sum"):
            f = getattr(_LIB, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_double] * 4
        _LIB.q128_dd_mul_relerr.restype = ctypes.c_double
        _LIB.q128_dd_mul_relerr.argtypes = [_vp, _vp, _vp]
        _LIB.q128_dd_add_abserr.restype = ctypes.c_double
        _LIB.q128_dd_add_abserr.argtypes = [_vp, _vp, _vp]
        _LIB.q128_from_dd.argtypes = [_vp, _c_i64, _vp]
        _LIB.q128_max_abs_diff.restype = ctypes.c_double
        _LIB.q128_max_abs_diff.argtypes = [_vp, _vp, _c_i64]
        _LIB.q128_max_abs.restype = ctypes.c_double
        _LIB.q128_max_abs.argtypes = [_vp, _c_i64]
        _LIB.port_force_scalar.argtypes = [ctypes.c_int]
        _LIB.port_set_order.argtypes = [ctypes.c_int]
        _LIB.port_get_order.restype = ctypes.c_int
        _LIB.port_max_threads.restype = ctypes.c_int
        _LIB.port_has_avx2.restype = ctypes.c_int
    return _LIB


def ref_lib():
    """The reference's own qgemm_/xgemm_ built from /root/reference, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libqxref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p)


ORDER_SIMD, ORDER_SCALAR = 0, 1


def set_order(order):
    """Operation order of the restated plain kernels: ORDER_SIMD (the reference's AVX2/FMA
    kernels) or ORDER_SCALAR (its kernels for hosts without AVX2/FMA).  Returns the previous
    order."""
    prev = lib().port_get_order()
    lib().port_set_order(int(order))
    return prev


def _ch(t):
    return ctypes.c_char(t.encode() if isinstance(t, str) else t)


def port_mgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nthreads=0):
    """CPU restatement of the reference mgemm_ (in place on C)."""
    rc = lib().port_mgemm(_ch(ta), _ch(tb), m, n, k, _p(alpha), _p(A), lda, _p(B), ldb,
                          _p(beta), _p(C), ldc, nthreads)
    if rc != 0:
        raise MemoryError("port_mgemm: pack buffer allocation failed")


def port_wgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nthreads=0):
    """CPU restatement of the reference wgemm_ (in place on C)."""
    rc = lib().port_wgemm(_ch(ta), _ch(tb), m, n, k, _p(alpha), _p(A), lda, _p(B), ldb,
                          _p(beta), _p(C), ldc, nthreads)
    if rc != 0:
        raise MemoryError("port_wgemm: pack buffer allocation failed")


def port_mgemmtr(uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb=8, nthreads=0):
    """CPU restatement of the reference mgemmtr_ (UPLO triangle of C, in place)."""
    rc = lib().port_mgemmtr(_ch(uplo), _ch(ta), _ch(tb), n, k, _p(alpha), _p(A), lda, _p(B),
                            ldb, _p(beta), _p(C), ldc, nb, nthreads)
    if rc != 0:
        raise MemoryError("port_mgemmtr: pack buffer allocation failed")


def port_wgemmtr(uplo, ta, tb, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb=8, nthreads=0):
    """CPU restatement of the reference wgemmtr_ (UPLO triangle of C, in place)."""
    rc = lib().port_wgemmtr(_ch(uplo), _ch(ta), _ch(tb), n, k, _p(alpha), _p(A), lda, _p(B),
                            ldb, _p(beta), _p(C), ldc, nb, nthreads)
    if rc != 0:
        raise MemoryError("port_wgemmtr: pack buffer allocation failed")


def port_mgemv(trans, m, n, alpha, A, lda, x, incx, beta, y, incy, nbands=1):
    """CPU restatement of the reference mgemv_ (in place on y).  nbands: the reference's
    OpenMP team size for op 'N' (row bands; 1 = a serial run)."""
    if lib().port_mgemv(_ch(trans), m, n, _p(alpha), _p(A), lda, _p(x), incx, _p(beta), _p(y),
                        incy, nbands) != 0:
        raise MemoryError("port_mgemv: scratch allocation failed")


def port_wgemv(trans, m, n, alpha, A, lda, x, incx, beta, y, incy):
    """CPU restatement of the reference wgemv_ (in place on y)."""
    if lib().port_wgemv(_ch(trans), m, n, _p(alpha), _p(A), lda, _p(x), incx, _p(beta), _p(y),
                        incy) != 0:
        raise MemoryError("port_wgemv: scratch allocation failed")


def _entries(ent):
    if ent is None:
        return -1, None, None, None
    ei = np.ascontiguousarray(ent[0], dtype=np.int64)
    ej = np.ascontiguousarray(ent[1], dtype=np.int64)
    return len(ei), ei, ej, (ei, ej)


def q128_gemm_err(cplx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, Ct, ldct,
                  entries=None):
    """Normwise error of Ct against the binary128 oracle: dict(err, dmax, rmax, opnorm).
    ``entries`` = (rows, cols) arrays selects a sample of output entries."""
    ne, ei, ej, keep = _entries(entries)
    out = np.zeros(4)
    f = lib().q128_wgemm_err if cplx else lib().q128_mgemm_err
    f(_ch(ta), _ch(tb), m, n, k, _p(alpha), _p(A), lda, _p(B), ldb, _p(beta), _p(C0), ldc,
      _p(Ct), ldct, ne, None if ei is None else _p(ei), None if ej is None else _p(ej),
      _p(out))
    del keep
    return {"err": out[0], "dmax": out[1], "rmax": out[2], "opnorm": out[3]}


def q128_gemm_dd(cplx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, entries=None):
    """Oracle result rounded to DD: shape (n, m, 2|4) for all entries, (nent, 2|4) for a
    sample; also returns the raw binary128 bytes."""
    ne, ei, ej, keep = _entries(entries)
    w = 4 if cplx else 2
    cnt = m * n if ne < 0 else ne
    out_dd = np.zeros((cnt, w))
    out_q = np.zeros(cnt * 8 * w, dtype=np.uint8)  # 16 bytes per binary128 component
    f = lib().q128_wgemm_entries if cplx else lib().q128_mgemm_entries
    f(_ch(ta), _ch(tb), m, n, k, _p(alpha), _p(A), lda, _p(B), ldb, _p(beta), _p(C0), ldc,
      ne, None if ei is None else _p(ei), None if ej is None else _p(ej), _p(out_q),
      _p(out_dd))
    del keep
    if ne < 0:
        out_dd = out_dd.reshape(n, m, w)
    return out_dd, out_q
