"""Seeded synthetic double-double operands in the distributions the workloads use.

  uniform : hi = s*(2u-1), u in [0,1) at 53-bit resolution;
            lo = hi * 2^-53 * (u' - 1/2)   (|lo| <= ulp(hi)/2: a normalised DD)
  wide    : hi = +-U[1,2) * 2^e, e uniform integer in [emin, emax]; lo as above
  integer : small integers (|x| <= bound) with lo = 0, for exact known-answer products

Padding rows (ld > rows) are filled with the reference fuzz sentinel
(-7.3737373737e18, 0) for real and (-7.37e18, 4.24e17) for complex entries
(test/epblas-parallel/common/target_multifloats/fuzz_util.fypp:43-47), so stray writes
and reads of padding are detectable.

Layout: shape (cols, ld, 2) for real, (cols, ld, 4) for complex (column-major).
"""
import numpy as np

SENTINEL_RE = -7.3737373737e18
SENTINEL_IM = 4.2424242424e17


def _u53(rng, shape):
    return (rng.integers(0, 1 << 53, size=shape, dtype=np.int64).astype(np.float64)
            * 2.0 ** -53)


def _with_lo(rng, hi):
    lo = hi * 2.0 ** -53 * (_u53(rng, hi.shape) - 0.5)
    return np.stack([hi, lo], axis=-1)


def real_values(rng, shape, kind="uniform", scale=1.0, emin=-20, emax=20, bound=64):
    """DD values of the given shape (+ trailing limb axis of 2)."""
    if kind == "uniform":
        hi = scale * (2.0 * _u53(rng, shape) - 1.0)
        return _with_lo(rng, hi)
    if kind == "wide":
        sign = np.where(rng.integers(0, 2, size=shape) == 0, 1.0, -1.0)
        e = rng.integers(emin, emax + 1, size=shape)
        hi = sign * np.ldexp(1.0 + _u53(rng, shape), e)
        return _with_lo(rng, hi)
    if kind == "integer":
        hi = rng.integers(-bound, bound + 1, size=shape).astype(np.float64)
        return np.stack([hi, np.zeros_like(hi)], axis=-1)
    if kind == "hi_only":   # the reference fuzz fill: lo limb = 0 (fuzz_util.fypp:216-232)
        hi = scale * (2.0 * _u53(rng, shape) - 1.0)
        return np.stack([hi, np.zeros_like(hi)], axis=-1)
    raise ValueError(kind)


def real_matrix(rng, rows, cols, ld=None, **kw):
    ld = rows if ld is None else ld
    out = np.empty((cols, ld, 2))
    out[..., 0] = SENTINEL_RE
    out[..., 1] = 0.0
    if rows > 0 and cols > 0:
        out[:, :rows, :] = real_values(rng, (cols, rows), **kw)
    return out


def cplx_matrix(rng, rows, cols, ld=None, **kw):
    ld = rows if ld is None else ld
    out = np.empty((cols, ld, 4))
    out[..., 0] = SENTINEL_RE
    out[..., 1] = 0.0
    out[..., 2] = SENTINEL_IM
    out[..., 3] = 0.0
    if rows > 0 and cols > 0:
        out[:, :rows, 0:2] = real_values(rng, (cols, rows), **kw)
        out[:, :rows, 2:4] = real_values(rng, (cols, rows), **kw)
    return out


def real_scalar(rng, kind="uniform", scale=1.0):
    return real_values(rng, (), kind=kind, scale=scale).reshape(2)


def cplx_scalar(rng, kind="uniform", scale=1.0):
    return np.concatenate([real_scalar(rng, kind, scale), real_scalar(rng, kind, scale)])


def fuzz_alpha_beta(rng, cplx=False):
    """alpha/beta as the reference fuzz draws them: 0, 1, -1 at 10% each, else U[-1,1]
    (fuzz_util.fypp:160-185)."""
    def one():
        r = rng.random()
        w = 4 if cplx else 2
        v = np.zeros(w)
        if r < 0.1:
            return v
        if r < 0.2:
            v[0] = 1.0
            return v
        if r < 0.3:
            v[0] = -1.0
            return v
        v[0] = 2.0 * rng.random() - 1.0
        if cplx:
            v[2] = 2.0 * rng.random() - 1.0
        return v
    return one(), one()


def rand_int_log(rng, lo, hi):
    """Log-uniform integer in [lo, hi] (fuzz_util.fypp:89-110)."""
    x = np.exp(rng.uniform(np.log(lo), np.log(hi + 1)))
    return int(min(hi, max(lo, int(x))))


def rand_incx(rng):
    """Vector stride as the reference fuzz draws it: 1 (50%), 2 (20%), 3 (15%), -1 (10%),
    -2 (5%) (fuzz_util.fypp:405-420)."""
    r = rng.random()
    return 1 if r < 0.5 else 2 if r < 0.7 else 3 if r < 0.85 else -1 if r < 0.95 else -2


def rand_pad(rng):
    """Leading-dimension padding 0-4 / 5-32 / 33-128 (fuzz_util.fypp:390-401)."""
    r = rng.random()
    if r < 0.5:
        return int(rng.integers(0, 5))
    if r < 0.85:
        return int(rng.integers(5, 33))
    return int(rng.integers(33, 129))
