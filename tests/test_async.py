"""Async regions (rts/dvmh/region.cpp:442: REGION_ASYNC honoured when
DVMH_ALLOW_ASYNC is set, rts/dvmh/settings.cpp:361).  programs/asyncreg.cdv
relaxes a grid in `region async` with SUM and MAX reductions and records, right
after each region, how many GPU reduction results are still pending; get_actual
then merges them.  Allowed, the loop and the region end without a
synchronisation (two results pending per step); not allowed, or on the host,
the results are merged at the loop's end (none pending).  Results: grid
bit-exact and MAX exact against a sequential evaluation, SUM within 1e-12
relative of math.fsum (the GPU folds in another order)."""
import ctypes
import math

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def reference(n, iters):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    A = ((i * 7 + j * 3) % 11).astype(np.float64) * 0.25
    sums, maxes = [], []
    for it in range(iters):
        A = A * 0.5 + ((i + j + it) % 5).astype(np.float64)
        sums.append(math.fsum(A.ravel()))
        maxes.append(A.max())
    return A.ravel(), sums, maxes


def run(n, iters, allow):
    rt = D.runtime()
    rt.dvmh_amd_set_allow_async_C.argtypes = [ctypes.c_long]
    rt.dvmh_amd_pending_posts_C.restype = ctypes.c_long
    f = D.load('libcdv_asyncreg.so').asyncreg
    f.argtypes = [ctypes.c_long, ctypes.c_long, _dp, _lp, _dp]
    sm = np.zeros(2 * iters)
    pend = np.full(iters, -1, dtype=np.int64)
    out = np.zeros(n * n)
    rt.dvmh_amd_set_allow_async_C(1 if allow else 0)
    try:
        f(n, iters, sm.ctypes.data_as(_dp), pend.ctypes.data_as(_lp), out.ctypes.data_as(_dp))
    finally:
        rt.dvmh_amd_set_allow_async_C(0)
    assert rt.dvmh_amd_pending_posts_C() == 0
    return sm, pend, out


def check(n, iters, sm, out):
    grid, sums, maxes = reference(n, iters)
    assert H.bits_equal(out, grid)
    for it in range(iters):
        assert abs(sm[2 * it] - sums[it]) <= 1e-12 * abs(sums[it]), (it, sm[2 * it], sums[it])
        assert sm[2 * it + 1] == maxes[it]


def test_async_region_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    sm, pend, out = run(9, 3, True)
    assert list(pend) == [0, 0, 0]
    check(9, 3, sm, out)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [5, 64, 257])
def test_async_region_defers_reductions(n):
    sm, pend, out = run(n, 4, True)
    assert list(pend) == [2, 2, 2, 2]
    check(n, 4, sm, out)


@pytest.mark.gpu
def test_async_region_not_allowed_synchronises():
    sm, pend, out = run(64, 3, False)
    assert list(pend) == [0, 0, 0]
    check(64, 3, sm, out)
