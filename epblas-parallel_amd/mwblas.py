"""Host-side mirror of the reference mwblas GEMM interface over libmwgemm_hip.so.

The functions keep the reference's names, argument order and meaning
(``mgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)``,
src/epblas-parallel/multifloats/mgemm_parallel.cpp:189-191 through
common/epblas_facade.h:234-240) and call the library's Fortran-ABI symbols by
reference, exactly as a Fortran or C caller of the reference would.

Array convention: a column-major DD matrix with leading dimension ``ld`` and ``cols``
columns is a float64 array of shape ``(cols, ld, 2)`` (C-contiguous) holding
``(hi, lo)``; complex DD uses shape ``(cols, ld, 4)`` holding
``(re.hi, re.lo, im.hi, im.lo)``.  Host numpy arrays and device ``torch`` tensors are
both accepted by the drop-in calls (device memory is used in place).

There is no CPU fallback: if the library is missing or a call fails, this module raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmwgemm_hip.so")

PATH_PLAIN, PATH_OZAKI, PATH_AUTO = 0, 1, 2

_LIB = None
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


class MwgemmError(RuntimeError):
    pass




def lib():
    """Load libmwgemm_hip.so (built by build.py); raise if it is not there."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MwgemmError(f"{LIB_PATH} is missing: run epblas-parallel_amd/build.py first")
    L = ctypes.CDLL(LIB_PATH)
    c = ctypes.c_char
    for name, ip in (("mgemm_", _i32p), ("mgemm_64_", _i64p), ("wgemm_", _i32p),
                     ("wgemm_64_", _i64p)):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ip, ip, ip, _vp, _vp, ip, _vp, ip, _vp,
                      _vp, ip]
    for name, ip in (("mgemmtr_", _i32p), ("mgemmtr_64_", _i64p), ("wgemmtr_", _i32p),
                     ("wgemmtr_64_", _i64p)):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ip, ip, _vp, _vp, ip,
                      _vp, ip, _vp, _vp, ip]
    for name in ("mgemm_serial", "wgemm_serial"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [c, c, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64]
    for name in ("mwgemm_mgemm_device", "mwgemm_wgemm_device"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [c, c, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                      ctypes.c_int, _vp]
    L.mwgemm_set_plain_order.restype = ctypes.c_int
    L.mwgemm_set_plain_order.argtypes = [ctypes.c_int]
    L.mwgemm_get_plain_order.restype = ctypes.c_int
    L.mwgemm_set_path.restype = ctypes.c_int
    L.mwgemm_set_path.argtypes = [ctypes.c_int]
    L.mwgemm_get_path.restype = ctypes.c_int
    L.mwgemm_last_error.restype = ctypes.c_int
    L.mwgemm_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.mwgemm_clear_error.restype = None
    L.mwgemm_set_abort.restype = ctypes.c_int
    L.mwgemm_set_abort.argtypes = [ctypes.c_int]
    L.mwgemm_timer_enable.restype = ctypes.c_int
    L.mwgemm_timer_enable.argtypes = [ctypes.c_int]
    L.mwgemm_timer_reset.restype = ctypes.c_int
    L.mwgemm_timer_read.restype = ctypes.c_int
    L.mwgemm_timer_read.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    if hasattr(L, "mwgemm_ozaki_config"):
        L.mwgemm_ozaki_config.restype = ctypes.c_int
        L.mwgemm_ozaki_config.argtypes = [ctypes.c_int64, _vp, _vp]
    _LIB = L
    # Errors come back as exceptions here rather than aborting the interpreter.
    L.mwgemm_set_abort(0)
    return L


def last_error():
    buf = ctypes.create_string_buffer(512)
    code = lib().mwgemm_last_error(buf, 512)
    return code, buf.value.decode(errors="replace")


def clear_error():
    lib().mwgemm_clear_error()


def _check(rc=None):
    code, msg = last_error()
    if (rc is not None and rc != 0) or code != 0:
        clear_error()
        raise MwgemmError(f"libmwgemm_hip error {code if code else rc}: {msg}")


def set_path(path):
    prev = lib().mwgemm_set_path(path)
    if prev < 0:
        raise ValueError(f"bad path {path}")
    return prev


ORDER_SIMD, ORDER_SCALAR = 0, 1


def set_plain_order(order):
    """Operation order of the plain path: ORDER_SIMD (the reference's AVX2/FMA kernels) or
    ORDER_SCALAR (its kernels for hosts without AVX2/FMA).  Returns the previous order."""
    prev = lib().mwgemm_set_plain_order(int(order))
    if prev < 0:
        raise ValueError(f"bad order {order}")
    return prev


def get_plain_order():
    return lib().mwgemm_get_plain_order()


def _ptr(x):
    """Data pointer of a numpy array or a torch tensor (host or device)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(x.data_ptr())


def _scalar(x, width):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    if a.size != width:
        raise ValueError(f"scalar must have {width} limbs")
    return a


def _gemm(name, width, ip, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    L = lib()
    ct = ctypes.c_int32 if ip is _i32p else ctypes.c_int64
    al, be = _scalar(alpha, width), _scalar(beta, width)
    args = [ct(v) for v in (m, n, k, lda, ldb, ldc)]
    clear_error()
    getattr(L, name)(transa.encode(), transb.encode(), ctypes.byref(args[0]),
                     ctypes.byref(args[1]), ctypes.byref(args[2]), _ptr(al), _ptr(a),
                     ctypes.byref(args[3]), _ptr(b), ctypes.byref(args[4]), _ptr(be), _ptr(c),
                     ctypes.byref(args[5]))
    _check()


def mgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """C := alpha*op(A)*op(B) + beta*C (real DD) through the LP64 drop-in ``mgemm_``."""
    _gemm("mgemm_", 2, _i32p, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def mgemm_64(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """ILP64 twin ``mgemm_64_``."""
    _gemm("mgemm_64_", 2, _i64p, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def wgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Complex DD GEMM through the LP64 drop-in ``wgemm_``."""
    _gemm("wgemm_", 4, _i32p, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def wgemm_64(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _gemm("wgemm_64_", 4, _i64p, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def mgemm_serial(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """The by-value core ``mgemm_serial`` (reference mgemm_kernel.h:68-75)."""
    al, be = _scalar(alpha, 2), _scalar(beta, 2)
    clear_error()
    lib().mgemm_serial(transa.encode(), transb.encode(), m, n, k, _ptr(al), _ptr(a), lda,
                       _ptr(b), ldb, _ptr(be), _ptr(c), ldc)
    _check()


def wgemm_serial(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    al, be = _scalar(alpha, 4), _scalar(beta, 4)
    clear_error()
    lib().wgemm_serial(transa.encode(), transb.encode(), m, n, k, _ptr(al), _ptr(a), lda,
                       _ptr(b), ldb, _ptr(be), _ptr(c), ldc)
    _check()


def _gemmtr(name, width, ip, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    L = lib()
    ct = ctypes.c_int32 if ip is _i32p else ctypes.c_int64
    al, be = _scalar(alpha, width), _scalar(beta, width)
    args = [ct(v) for v in (n, k, lda, ldb, ldc)]
    clear_error()
    getattr(L, name)(uplo.encode(), transa.encode(), transb.encode(), ctypes.byref(args[0]),
                     ctypes.byref(args[1]), _ptr(al), _ptr(a), ctypes.byref(args[2]), _ptr(b),
                     ctypes.byref(args[3]), _ptr(be), _ptr(c), ctypes.byref(args[4]))
    _check()


def mgemmtr(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """C := alpha*op(A)*op(B) + beta*C on the UPLO triangle of the n x n C (drop-in
    ``mgemmtr_``)."""
    _gemmtr("mgemmtr_", 2, _i32p, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def mgemmtr_64(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _gemmtr("mgemmtr_64_", 2, _i64p, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
            ldc)


def wgemmtr(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Complex triangular-output GEMM (drop-in ``wgemmtr_``)."""
    _gemmtr("wgemmtr_", 4, _i32p, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def wgemmtr_64(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _gemmtr("wgemmtr_64_", 4, _i64p, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
            ldc)


def gemmtr_serial(cplx, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """The by-value cores ``mgemmtr_serial`` / ``wgemmtr_serial``."""
    w = 4 if cplx else 2
    al, be = _scalar(alpha, w), _scalar(beta, w)
    f = lib().wgemmtr_serial if cplx else lib().mgemmtr_serial
    f.restype = None
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int64, ctypes.c_int64,
                  _vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, _vp, _vp, ctypes.c_int64]
    clear_error()
    f(uplo.encode(), transa.encode(), transb.encode(), n, k, _ptr(al), _ptr(a), lda, _ptr(b),
      ldb, _ptr(be), _ptr(c), ldc)
    _check()


def gemmtr_device(cplx, uplo, transa, transb, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc,
                  stream=None):
    """Device-pointer triangular-output GEMM (asynchronous on ``stream``)."""
    L = lib()
    w = 4 if cplx else 2
    al, be = _scalar(alpha, w), _scalar(beta, w)
    f = L.mwgemm_wgemmtr_device if cplx else L.mwgemm_mgemmtr_device
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int64, ctypes.c_int64,
                  _vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, _vp, _vp, ctypes.c_int64, _vp]
    clear_error()
    rc = f(uplo.encode(), transa.encode(), transb.encode(), n, k, _ptr(al), _ptr(dA), lda,
           _ptr(dB), ldb, _ptr(be), _ptr(dC), ldc,
           None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


def _gemv(name, width, ct, trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    L = lib()
    al, be = _scalar(alpha, width), _scalar(beta, width)
    args = [ct(v) for v in (m, n, lda, incx, incy)]
    f = getattr(L, name)
    f.restype = None
    ip = ctypes.POINTER(ct)
    f.argtypes = [ctypes.c_char_p, ip, ip, _vp, _vp, ip, _vp, ip, _vp, _vp, ip]
    clear_error()
    f(trans.encode(), ctypes.byref(args[0]), ctypes.byref(args[1]), _ptr(al), _ptr(a),
      ctypes.byref(args[2]), _ptr(x), ctypes.byref(args[3]), _ptr(be), _ptr(y),
      ctypes.byref(args[4]))
    _check()


def mgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    """y := alpha*op(A)*x + beta*y (real DD) through the drop-in ``mgemv_``.  Vectors are
    (span, 2) arrays holding every |inc|-th entry."""
    _gemv("mgemv_", 2, ctypes.c_int32, trans, m, n, alpha, a, lda, x, incx, beta, y, incy)


def mgemv_64(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    _gemv("mgemv_64_", 2, ctypes.c_int64, trans, m, n, alpha, a, lda, x, incx, beta, y, incy)


def wgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    """Complex DD matrix-vector product through the drop-in ``wgemv_``."""
    _gemv("wgemv_", 4, ctypes.c_int32, trans, m, n, alpha, a, lda, x, incx, beta, y, incy)


def wgemv_64(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    _gemv("wgemv_64_", 4, ctypes.c_int64, trans, m, n, alpha, a, lda, x, incx, beta, y, incy)


def gemv_device(cplx, trans, m, n, alpha, dA, lda, dx, incx, beta, dy, incy, nbands=0,
                stream=None):
    """Device-pointer matrix-vector product (asynchronous on ``stream``)."""
    L = lib()
    w = 4 if cplx else 2
    al, be = _scalar(alpha, w), _scalar(beta, w)
    i64 = ctypes.c_int64
    if cplx:
        f = L.mwgemm_wgemv_device
        f.argtypes = [ctypes.c_char, i64, i64, _vp, _vp, i64, _vp, i64, _vp, _vp, i64, _vp]
        extra = ()
    else:
        f = L.mwgemm_mgemv_device
        f.argtypes = [ctypes.c_char, i64, i64, _vp, _vp, i64, _vp, i64, _vp, _vp, i64, i64, _vp]
        extra = (nbands,)
    f.restype = ctypes.c_int
    clear_error()
    rc = f(trans.encode(), m, n, _ptr(al), _ptr(dA), lda, _ptr(dx), incx, _ptr(be), _ptr(dy),
           incy, *extra, None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


def gemm_device(cplx, transa, transb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc,
                path=PATH_PLAIN, stream=None):
    """Device-pointer GEMM (tensors on the GPU), asynchronous on ``stream`` (an int handle
    such as ``torch.cuda.current_stream().cuda_stream``; None or 0 = HIP's default stream,
    which is torch's default stream)."""
    L = lib()
    w = 4 if cplx else 2
    al, be = _scalar(alpha, w), _scalar(beta, w)
    f = L.mwgemm_wgemm_device if cplx else L.mwgemm_mgemm_device
    clear_error()
    rc = f(transa.encode(), transb.encode(), m, n, k, _ptr(al), _ptr(dA), lda, _ptr(dB), ldb,
           _ptr(be), _ptr(dC), ldc, path, None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


def gemm_device_noalloc(transa, transb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc,
                        path=PATH_PLAIN, stream=None):
    """Real device-pointer GEMM through mwgemm_mgemm_device_noalloc: the same result as
    gemm_device, with the exact-product workspace in the library's static scratch memory."""
    f = lib().mwgemm_mgemm_device_noalloc
    f.restype = ctypes.c_int
    i64 = ctypes.c_int64
    f.argtypes = [ctypes.c_char, ctypes.c_char, i64, i64, i64, _vp, _vp, i64, _vp, i64, _vp, _vp,
                  i64, ctypes.c_int, _vp]
    al, be = _scalar(alpha, 2), _scalar(beta, 2)
    clear_error()
    rc = f(transa.encode(), transb.encode(), m, n, k, _ptr(al), _ptr(dA), lda, _ptr(dB), ldb,
           _ptr(be), _ptr(dC), ldc, path, None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


def timer_enable(on=True):
    return lib().mwgemm_timer_enable(1 if on else 0)


def timer_reset():
    lib().mwgemm_timer_reset()


def timer_read():
    """{kernel class: (total ms, launches)} since the last reset (synchronises)."""
    cap = 32
    ms = np.zeros(cap)
    cnt = np.zeros(cap, dtype=np.int64)
    names = ctypes.create_string_buffer(2048)
    n = lib().mwgemm_timer_read(_ptr(ms), _ptr(cnt), cap, names, 2048)
    labels = names.raw.split(b"\0")[:n]
    return {labels[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(n) if cnt[i]}


def ozaki_config(k, cplx=False):
    """(number of moduli, integer width b) the exact-product path uses at depth k."""
    L = lib()
    f = L.mwgemm_ozaki_config
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64, ctypes.c_int, _vp, _vp]
    nm, b = ctypes.c_int(), ctypes.c_int()
    if f(k, 1 if cplx else 0, ctypes.byref(nm), ctypes.byref(b)) != 0:
        raise MwgemmError(f"no Ozaki plan for k={k}")
    return nm.value, b.value


def ozaki_set(bits=0, nmod=0):
    """Integer width (0 = automatic: 102 with the data-dependent modulus count, else at
    least 103) and forced modulus count (0 = automatic)."""
    f = lib().mwgemm_ozaki_set
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int]
    if f(bits, nmod) != 0:
        raise ValueError("bad Ozaki parameters")


ESTIMATE_OFF, ESTIMATE_LARGE, ESTIMATE_ALWAYS = 0, 1, 2


def ozaki_set_estimate(mode):
    """Data-dependent modulus count of real exact products: ESTIMATE_OFF, ESTIMATE_LARGE
    (products of at least 2^33 multiply-adds; the default) or ESTIMATE_ALWAYS."""
    f = lib().mwgemm_ozaki_set_estimate
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int]
    if f(int(mode)) != 0:
        raise ValueError(f"bad estimate mode {mode}")


def testing_ozaki_count_offset(offset):
    """Test hook: offset added to the data-dependent modulus count (0 = normal)."""
    f = lib().mwgemm_testing_ozaki_count_offset
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int]
    f(offset)


def ozaki_last_moduli():
    """The modulus count the last real exact product on the current device kept with the
    data-dependent count (waits for the device; 0 when none has run)."""
    f = lib().mwgemm_ozaki_last_moduli
    f.restype = ctypes.c_int
    f.argtypes = [_vp]
    nm = ctypes.c_int()
    if f(ctypes.byref(nm)) != 0:
        raise MwgemmError("mwgemm_ozaki_last_moduli failed")
    return nm.value


def ozaki_moduli(nmod):
    """The first nmod moduli and their CRT weights y_t for that set."""
    f = lib().mwgemm_ozaki_moduli
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, _vp, _vp]
    m = np.zeros(nmod, dtype=np.int32)
    y = np.zeros(nmod, dtype=np.int32)
    if f(nmod, _ptr(m), _ptr(y)) != 0:
        raise ValueError("bad modulus count")
    return m, y


def testing_residue_gemm(dA, dB, dU, mp, np_, kp, mod_index, variant=1, stream=None):
    """Run the production INT8 residue GEMM kernel of one tile variant once (test hook):
    0 = 256x256 DMA ring, 1 = 128x128, 2 = 256x128, 3 = 128x256."""
    f = lib().mwgemm_testing_residue_gemm
    f.restype = ctypes.c_int
    f.argtypes = [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int, _vp]
    rc = f(_ptr(dA), _ptr(dB), _ptr(dU), mp, np_, kp, mod_index, variant,
           None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


def residue_products(U, mp, np_):
    """The reduced residue products written by testing_residue_gemm as an (np_, mp) int8
    array (column j, row i): symmetric residues, |u| <= (m - 1) / 2 (+-128 for m = 256).
    The library stores them in 32 x 32 blocks in MFMA accumulator order: blocks column block
    by column block (ldu = mp), within a block word g*64 + l (g = 0..3, lane l = 0..63)
    packing rows 8g + 4(l >> 5) .. +3 of column l & 31, one byte per row."""
    import numpy as np
    u = np.asarray(U).reshape(-1)
    if u.dtype == np.uint8:
        u = u.view(np.int8)       # the bytes are signed residues
    assert mp % 32 == 0 and np_ % 32 == 0 and u.size >= mp * np_
    blk = u[:mp * np_].reshape(np_ // 32, mp // 32, 4, 2, 32, 4)   # bj, bi, g, half, col, q
    return blk.transpose(0, 4, 1, 2, 3, 5).reshape(np_, mp)


def testing_set_alloc_limit(nbytes):
    """Test hook: every device buffer the library would allocate for itself above nbytes
    counts as unobtainable (0 lifts the limit).  Returns the limit that was set."""
    f = lib().mwgemm_testing_set_alloc_limit
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64]
    if f(int(nbytes)) != 0:
        raise ValueError("bad allocation limit")
    return int(nbytes)


def _batched_args(ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs, ldcs):
    cnt = len(ms)
    i64 = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.int64).reshape(cnt))
    f2 = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(cnt, 2))
    ptrs = lambda xs: np.ascontiguousarray(
        np.array([0 if x is None else x.data_ptr() for x in xs], dtype=np.uint64))
    return cnt, [i64(ms), i64(ns), i64(ks), f2(alphas), ptrs(As), i64(ldas), ptrs(Bs), i64(ldbs),
                 f2(betas), ptrs(Cs), i64(ldcs)]


def _batched_sig(f, with_plan):
    f.restype = ctypes.c_int
    head = [_vp] if with_plan else []
    f.argtypes = head + [ctypes.c_int64, ctypes.c_char, ctypes.c_char] + [_vp] * 11 + [ctypes.c_int] + ([] if with_plan else [_vp])


class BatchedPlan:
    """A captured batched DD GEMM (see mwgemm_batched_plan_create).  Matrices are device
    tensors in the (cols, ld, 2) layout; the plan is bound to them."""

    def __init__(self, transa, transb, ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs, ldcs,
                 path=PATH_PLAIN):
        L = lib()
        _batched_sig(L.mwgemm_batched_plan_create, True)
        cnt, arrs = _batched_args(ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs, ldcs)
        self._keep = arrs
        self.handle = ctypes.c_void_p()
        clear_error()
        rc = L.mwgemm_batched_plan_create(ctypes.byref(self.handle), cnt, transa.encode(),
                                          transb.encode(), *[_ptr(a) for a in arrs], path)
        if rc != 0:
            _check(rc)

    def run(self, stream=None):
        f = lib().mwgemm_batched_plan_run
        f.restype = ctypes.c_int
        f.argtypes = [_vp, _vp]
        rc = f(self.handle, None if stream is None else ctypes.c_void_p(stream))
        if rc != 0:
            _check(rc)

    def info(self):
        f = lib().mwgemm_batched_plan_info
        f.restype = ctypes.c_int
        f.argtypes = [_vp, _vp, _vp, _vp, _vp]
        p, nm = ctypes.c_int(), ctypes.c_int()
        t, fl = ctypes.c_int64(), ctypes.c_double()
        f(self.handle, ctypes.byref(p), ctypes.byref(t), ctypes.byref(fl), ctypes.byref(nm))
        return {"path": p.value, "tiles": t.value, "flops": fl.value, "moduli": nm.value}

    def close(self):
        if self.handle:
            f = lib().mwgemm_batched_plan_destroy
            f.restype = ctypes.c_int
            f.argtypes = [_vp]
            f(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mgemm_batched_device(transa, transb, ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs,
                         ldcs, path=PATH_PLAIN, stream=None):
    """Build, run (no graph) and free a batched plan; synchronises the stream."""
    L = lib()
    _batched_sig(L.mwgemm_mgemm_batched_device, False)
    cnt, arrs = _batched_args(ms, ns, ks, alphas, As, ldas, Bs, ldbs, betas, Cs, ldcs)
    clear_error()
    rc = L.mwgemm_mgemm_batched_device(cnt, transa.encode(), transb.encode(),
                                       *[_ptr(a) for a in arrs], path,
                                       None if stream is None else ctypes.c_void_p(stream))
    if rc != 0:
        _check(rc)


TUNE_PLAIN_REAL, TUNE_PLAIN_CPLX, TUNE_OZ_GEMM, TUNE_BATCH = 0, 1, 2, 3


def autotune(m, n, k, cplx=False, path=PATH_PLAIN, all_times=False):
    """Time the tile variants of a path on synthetic m x n x k operands; returns
    (winning variant, its milliseconds[, every variant's milliseconds]) and records the
    decision."""
    f = lib().mwgemm_autotune
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                  _vp, _vp, _vp, ctypes.c_int]
    v, ms = ctypes.c_int(), ctypes.c_double()
    allv = np.full(8, -1.0)
    clear_error()
    rc = f(m, n, k, 1 if cplx else 0, path, ctypes.byref(v), ctypes.byref(ms), _ptr(allv), 8)
    if rc != 0:
        _check(rc)
    if all_times:
        return v.value, ms.value, [float(x) for x in allv if x >= 0]
    return v.value, ms.value


def batched_autotune(transa, transb, ms, ns, ks):
    """Tune a batch class (plain vs Ozaki residue tiles); returns (choice, milliseconds) with
    choice = path * 16 + residue-tile variant."""
    f = lib().mwgemm_batched_autotune
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64, ctypes.c_char, ctypes.c_char, _vp, _vp, _vp, _vp, _vp]
    cnt = len(ms)
    arr = [np.ascontiguousarray(np.asarray(x, dtype=np.int64)) for x in (ms, ns, ks)]
    c, t = ctypes.c_int(), ctypes.c_double()
    clear_error()
    rc = f(cnt, transa.encode(), transb.encode(), *[_ptr(a) for a in arr], ctypes.byref(c),
           ctypes.byref(t))
    if rc != 0:
        _check(rc)
    return c.value, t.value


def autotune_lookup(kind, a, b, c):
    f = lib().mwgemm_autotune_lookup
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
    return f(kind, a, b, c)


def autotune_clear():
    lib().mwgemm_autotune_clear()
