"""2-D Jacobi (plain method): the converted program through libdvmh_amd,
host handlers on the CPU and HIP handlers on the GPU, against the oracle."""
import numpy as np
import pytest

import dvmh_amd as D
import helpers as H


@pytest.mark.parametrize('n,iters', [(5, 3), (37, 9), (130, 4)])
def test_host_path_bit_exact(n, iters):
    res = H.run_jac2d(n, iters, targets=D.DEVICE_TYPE_HOST)
    ref = H.oracle_jac2d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])


def _gpu_run(n, iters, **kw):
    before = D.stats()
    res = H.run_jac2d(n, iters, **kw)
    after = D.stats()
    assert after['gpu_present'] == 1
    assert after['loops_gpu'] - before['loops_gpu'] == 1 + 2 * iters
    assert after['loops_host'] == before['loops_host']
    return res


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(3, 2), (5, 3), (37, 9), (130, 4), (257, 7), (1000, 3)])
def test_gpu_bit_exact(n, iters):
    res = _gpu_run(n, iters)
    ref = H.oracle_jac2d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
def test_gpu_full_size_properties():
    """2048^2 x 100 iterations (the workload's size): eps decreases monotonically
    (Jacobi is a contraction in max-norm for this problem), the boundary stays
    1.0 exactly and the first and last iterations match the oracle."""
    n, iters = 2048, 100
    res = _gpu_run(n, iters)
    A = res['A'].reshape(n, n)
    assert np.all(A[0, :] == 1.0) and np.all(A[-1, :] == 1.0) and np.all(A[:, 0] == 1.0) and np.all(A[:, -1] == 1.0)
    assert np.all(np.diff(res['eps']) <= 0)
    ref = H.oracle_jac2d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['eps'], ref['eps'])
