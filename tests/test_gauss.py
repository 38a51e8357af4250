"""The constructs of the reference's Gaussian-elimination examples
(examples/gauss_c.cdv, examples/gauss_wb.cdv) in programs/gauss.cdv, converted
by the front end: arrays declared without a distribution, redistribute with
WGTBLOCK weights (rts/dvmh/dvmh2.cpp dvmh_redistribute_C), realign
(dvmh_realign_C), remote_access in a loop clause and on a host block, a region
over sub-arrays, and statements of a region outside parallel loops (sequential
parts, cdvmh/src/converter_main.cpp:1310-1330, run where the region runs).
The eliminated matrix and the solution are compared bit for bit with a numpy
restatement of the same operation order; the restatement itself is pinned by
the solution solving the original system (numpy.linalg.solve)."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
from test_cdvmh import _u01

_dp = ctypes.POINTER(ctypes.c_double)


def reference(n, seed):
    A = _u01(seed, np.arange(n * (n + 1), dtype=np.uint64)).reshape(n, n + 1)
    A[np.arange(n), np.arange(n)] += float(n)
    A0 = A.copy()
    for i in range(n - 1):
        A[i + 1:, i + 1:] = A[i + 1:, i + 1:] - A[i + 1:, i:i + 1] * A[i, i + 1:] / A[i, i]
    X = np.zeros(n)
    X[n - 1] = A[n - 1, n] / A[n - 1, n - 1]
    for j in range(n - 2, -1, -1):
        A[:j + 1, n] = A[:j + 1, n] - A[:j + 1, j + 1] * X[j + 1]
        X[j] = A[j, n] / A[j, j]
    return A.ravel(), X, A0


def run(n, seed):
    f = D.load('libcdv_gauss.so').gauss
    f.argtypes = [ctypes.c_long, ctypes.c_ulonglong, _dp, _dp, _dp]
    wb = np.arange(1, n + 1, dtype=np.float64)
    x = np.full(n, np.nan)
    out = np.zeros(n * (n + 1))
    f(n, seed, wb.ctypes.data_as(_dp), x.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out, x


def check(n, seed=D.SEED_BASE + 40):
    out, x = run(n, seed)
    A, X, A0 = reference(n, seed)
    assert np.array_equal(out.view(np.uint64), A.view(np.uint64))
    assert np.array_equal(x.view(np.uint64), X.view(np.uint64))
    assert np.allclose(X, np.linalg.solve(A0[:, :n], A0[:, n]), rtol=1e-10, atol=0)


def test_restatement_solves_the_system():
    for n in (1, 2, 5, 40):
        A, X, A0 = reference(n, 7)
        assert np.allclose(A0[:, :n] @ X, A0[:, n], rtol=1e-12, atol=1e-12)


def test_gauss_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    for n in (1, 2, 9):
        check(n)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [1, 2, 7, 100, 300])
def test_gauss_gpu(n):
    before = D.stats()['loops_gpu']
    check(n)
    assert D.stats()['loops_gpu'] > before


def seqpart_reference(n, steps):
    i, j = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing='ij')
    A = i + 0.5 * j
    ts = []
    for it in range(steps):
        t = A[it % n, (it * 3) % n] * 2.0 + float(it)
        A = A * 0.5 + t
        ts.append(t)
    return np.array(ts), A.ravel()


def run_seqpart(n, steps):
    f = D.load('libcdv_seqpart.so').seqpart
    f.argtypes = [ctypes.c_long, ctypes.c_long, _dp, _dp]
    ts = np.full(steps, np.nan)
    out = np.zeros(n * n)
    f(n, steps, ts.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return ts, out


def check_seqpart(n, steps):
    ts, out = run_seqpart(n, steps)
    rts, rout = seqpart_reference(n, steps)
    assert np.array_equal(ts.view(np.uint64), rts.view(np.uint64))
    assert np.array_equal(out.view(np.uint64), rout.view(np.uint64))


def test_seqpart_host_path():
    """programs/seqpart.cdv: a sequential part writes the region's scalar,
    which the parallel loop reads and the host reads after the region."""
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    check_seqpart(7, 4)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [3, 64, 257])
def test_seqpart_gpu(n):
    before = D.stats()['loops_gpu']
    check_seqpart(n, 5)
    assert D.stats()['loops_gpu'] > before
