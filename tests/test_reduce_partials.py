"""dvmh_amd_reduce_partials_C: the runtime's device fold of a reduction's
partial values (the finish of every GPU loop reduction), called directly on
device buffers.  Checked against numpy on the same seeded partials: MAX / MIN
exact, integer SUM exact, FP64 SUM within 1e-12 relative of math.fsum,
MAXLOC / MINLOC value exact and loc the lowest among equal values (the
sequential strict-compare winner); sizes cover the one-pass fold and the
two-pass fold through scratch (more than 16384 partials)."""
import ctypes
import math

import numpy as np
import pytest

import dvmh_amd as D

RF_SUM, RF_MAX, RF_MIN, RF_MAXLOC, RF_MINLOC = 1, 3, 4, 11, 12
RT_LONG, RT_DOUBLE = 2, 4   # include/dvmhlib_const.h
SIZES = [1, 7, 1024, 16384, 16385, 100003, 3000000]


class Hip:
    """The few HIP runtime calls the test needs, through ctypes."""

    def __init__(self):
        h = ctypes.CDLL('libamdhip64.so')
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h = h

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def _rt():
    r = D.runtime()
    f = r.dvmh_amd_reduce_partials_C
    f.argtypes = [ctypes.c_long] * 3 + [ctypes.c_void_p] * 2 + [ctypes.c_long] + [ctypes.c_void_p] * 2
    f.restype = ctypes.c_long
    return f


def partials(n, seed, dtype):
    """Seeded partials with many repeated values (ties for MAX/MIN) and loc
    records that are a permutation of positions (so the lowest loc is not
    simply the first partial)."""
    rng = np.random.default_rng(seed)
    if dtype == np.float64:
        vals = rng.integers(0, max(2, n // 3), n).astype(np.float64) * 0.125 - 3.0
    else:
        vals = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    locs = rng.permutation(n).astype(np.int64) * 3 + 7
    return vals, locs


def expected(func, vals, locs):
    if func == RF_SUM:
        return (math.fsum(vals) if vals.dtype == np.float64 else int(np.sum(vals))), None
    if func == RF_MAX:
        return vals.max(), None
    if func == RF_MIN:
        return vals.min(), None
    v = vals.max() if func == RF_MAXLOC else vals.min()
    return v, int(locs[vals == v].min())


def run(f, hip, func, rt_type, vals, locs):
    dv = hip.upload(vals)
    dl = hip.upload(locs)
    out = hip.alloc(64)
    outl = hip.alloc(64)
    try:
        assert f(func, rt_type, len(vals), dv, dl, 8, out, outl) == 0
        v = hip.download(out, vals[:1])[0]
        loc = int(hip.download(outl, locs[:1])[0])
    finally:
        hip.free(dv, dl, out, outl)
    return v, loc


def test_reduce_partials_needs_gpu():
    if D.stats()['gpu_present']:
        pytest.skip('checked on the GPU by the tests below')
    v = np.zeros(4)
    assert _rt()(RF_MAX, RT_DOUBLE, 4, v.ctypes.data, None, 0, v.ctypes.data, None) == -1


@pytest.mark.gpu
@pytest.mark.parametrize('n', SIZES)
@pytest.mark.parametrize('func', [RF_SUM, RF_MAX, RF_MIN, RF_MAXLOC, RF_MINLOC])
def test_reduce_partials_double(func, n):
    f, hip = _rt(), Hip()
    vals, locs = partials(n, 1000 + n + func, np.float64)
    v, loc = run(f, hip, func, RT_DOUBLE, vals, locs)
    want, wloc = expected(func, vals, locs)
    if func == RF_SUM:
        assert abs(v - want) <= 1e-12 * max(abs(want), 1.0), (v, want)
    else:
        assert v == want
    if wloc is not None:
        assert loc == wloc, (loc, wloc)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [1, 16385, 3000000])
@pytest.mark.parametrize('func', [RF_SUM, RF_MAX, RF_MINLOC])
def test_reduce_partials_long(func, n):
    f, hip = _rt(), Hip()
    vals, locs = partials(n, 2000 + n + func, np.int64)
    v, loc = run(f, hip, func, RT_LONG, vals, locs)
    want, wloc = expected(func, vals, locs)
    assert int(v) == want
    if wloc is not None:
        assert loc == wloc
