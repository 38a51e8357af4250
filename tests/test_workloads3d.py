"""3-D workloads through libdvmh_amd against the oracle:
3-D 7-point Jacobi (MAX + SUM), lexicographic SOR (ACROSS on every axis,
wavefront handlers) and ADI line recurrences (ACROSS on one axis, MAXLOC)."""
import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

HOST = D.DEVICE_TYPE_HOST


def _on_gpu(fn, *a, **kw):
    before = D.stats()
    res = fn(*a, **kw)
    after = D.stats()
    assert after['gpu_present'] == 1
    assert after['loops_gpu'] > before['loops_gpu']
    assert after['loops_host'] == before['loops_host'], 'a loop ran on the host'
    return res


# ---------------------------------------------------------------- host paths (CPU)
@pytest.mark.parametrize('n,iters', [(4, 3), (19, 5)])
def test_jac3d_host(n, iters):
    r = H.run_jac3d(n, iters, targets=HOST)
    o = H.oracle_jac3d(n, iters)
    assert H.bits_equal(r['A'], o['A']) and H.bits_equal(r['B'], o['B'])
    assert np.array_equal(r['eps'], o['eps'])
    assert H.sum_close(r['sum'], o['sum'])


@pytest.mark.parametrize('n,iters', [(4, 2), (17, 4)])
def test_sor3d_host(n, iters):
    r = H.run_sor3d(n, iters, targets=HOST)
    o = H.oracle_sor3d(n, iters)
    assert H.bits_equal(r['A'], o['A'])
    assert np.array_equal(r['eps'], o['eps'])


@pytest.mark.parametrize('n,steps', [(3, 2), (13, 3)])
def test_adi3d_host(n, steps):
    r = H.run_adi3d(n, steps, targets=HOST)
    o = H.oracle_adi3d(n, steps)
    assert H.bits_equal(r['A'], o['A'])
    assert np.array_equal(r['max'], o['max'])
    assert np.array_equal(r['loc'], o['loc'])


# ---------------------------------------------------------------- GPU parity
@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(3, 2), (4, 3), (19, 5), (66, 4), (131, 3)])
def test_jac3d_gpu(n, iters):
    r = _on_gpu(H.run_jac3d, n, iters)
    o = H.oracle_jac3d(n, iters)
    assert H.bits_equal(r['A'], o['A']) and H.bits_equal(r['B'], o['B'])
    assert np.array_equal(r['eps'], o['eps'])
    assert H.sum_close(r['sum'], o['sum'])


@pytest.mark.gpu
def test_jac3d_gpu_full_size():
    """384^3 (the workload's grid) for 4 iterations: grid bit-exact, MAX exact,
    SUM within 1e-12 relative of the compensated oracle sum."""
    n, iters = 384, 4
    r = _on_gpu(H.run_jac3d, n, iters)
    o = H.oracle_jac3d(n, iters)
    assert H.bits_equal(r['A'], o['A']) and H.bits_equal(r['B'], o['B'])
    assert np.array_equal(r['eps'], o['eps'])
    assert H.sum_close(r['sum'], o['sum'])


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(3, 1), (4, 2), (9, 3), (17, 4), (40, 2)])
def test_sor3d_gpu(n, iters):
    r = _on_gpu(H.run_sor3d, n, iters)
    o = H.oracle_sor3d(n, iters)
    assert H.bits_equal(r['A'], o['A'])
    assert np.array_equal(r['eps'], o['eps'])


@pytest.mark.gpu
@pytest.mark.parametrize('n,steps', [(2, 2), (3, 2), (13, 3), (64, 2), (100, 1)])
def test_adi3d_gpu(n, steps):
    r = _on_gpu(H.run_adi3d, n, steps)
    o = H.oracle_adi3d(n, steps)
    assert H.bits_equal(r['A'], o['A'])
    assert np.array_equal(r['max'], o['max'])
    assert np.array_equal(r['loc'], o['loc'])
