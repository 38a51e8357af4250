"""dvmh_amd_reduce_partials_noalloc_C against dvmh_amd_reduce_partials_C:
the same deterministic fold, so value and loc must be identical bit for bit.
Two calls in a row share the fixed scratch: the first folds partials in two
passes (through the scratch), the second a smaller input, also through it;
a size above the scratch's limit is rejected without any device work."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
from test_reduce_partials import RF_MAX, RF_MAXLOC, RF_MINLOC, RF_SUM, RT_DOUBLE, RT_LONG, Hip, partials

MAX_PARTIALS = 715784192   # 1 MiB of scratch: 24 bytes per 16384 partials plus 64


def _fns():
    r = D.runtime()
    out = []
    for name in ('dvmh_amd_reduce_partials_C', 'dvmh_amd_reduce_partials_noalloc_C'):
        f = getattr(r, name)
        f.argtypes = [ctypes.c_long] * 3 + [ctypes.c_void_p] * 2 + [ctypes.c_long] + [ctypes.c_void_p] * 2
        f.restype = ctypes.c_long
        out.append(f)
    return out


def _both(hip, func, rt_type, vals, locs):
    ref, noalloc = _fns()
    dv = hip.upload(vals)
    dl = hip.upload(locs)
    outs = [hip.alloc(64) for _ in range(4)]
    try:
        assert ref(func, rt_type, len(vals), dv, dl, 8, outs[0], outs[1]) == 0
        assert noalloc(func, rt_type, len(vals), dv, dl, 8, outs[2], outs[3]) == 0
        got = [hip.download(o, locs[:1]).view(np.uint64)[0] for o in outs]
    finally:
        hip.free(dv, dl, *outs)
    return got


@pytest.mark.gpu
@pytest.mark.parametrize('func,rt_type', [(RF_SUM, RT_DOUBLE), (RF_MAX, RT_DOUBLE), (RF_MAXLOC, RT_DOUBLE),
                                          (RF_MINLOC, RT_LONG), (RF_SUM, RT_LONG)])
def test_no_alloc_fold_equals_allocating_fold(func, rt_type):
    hip = Hip()
    dtype = np.float64 if rt_type == RT_DOUBLE else np.int64
    for n in (2000003, 40001):   # second call smaller, both through the scratch
        print('no-alloc fold: func', func, 'type', rt_type, 'n', n, flush=True)
        vals, locs = partials(n, 77 + n + func, dtype)
        a_val, a_loc, b_val, b_loc = _both(hip, func, rt_type, vals, locs)
        assert a_val == b_val
        if func in (RF_MAXLOC, RF_MINLOC):
            assert a_loc == b_loc


@pytest.mark.gpu
def test_no_alloc_rejects_oversize():
    """Beyond the scratch's limit: -2 and nothing launched (the arguments are
    never dereferenced, so a one-element buffer stands for the partials)."""
    hip = Hip()
    _, noalloc = _fns()
    v = np.zeros(2)
    dv = hip.upload(v)
    out = hip.alloc(64)
    try:
        assert noalloc(RF_MAX, RT_DOUBLE, MAX_PARTIALS + 1, dv, None, 0, out, None) == -2
        assert noalloc(RF_MAX, RT_DOUBLE, 2, dv, None, 0, out, None) == 0
        assert hip.download(out, v[:1])[0] == 0.0
    finally:
        hip.free(dv, out)
