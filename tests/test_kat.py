"""The reference's own known-answer programs (examples/Adi/adi.c,
examples/Performance/jac3d_perf.c, jac2d_perf.c): their built-in check values
pin the oracle on the CPU and the whole GPU path end to end."""
import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

ADI_GOLD = (0.07249074, 1e-6)      # examples/Adi/adi.c:79, n = 384, 100 iterations
JAC3_GOLD = (5.058044, 1e-4)       # examples/Performance/jac3d_perf.c:92, L = 384, 100 iterations
JAC2_GOLD = (58.37598, 1e-3)       # examples/Performance/jac2d_perf.c:89, L = 8000, 100 iterations


# ---------------------------------------------------------------- oracle pinned by the golden values
def test_oracle_adi_kat_golden():
    eps = H.oracle().orc_adi_kat(384, 100, 0, None)
    assert abs(eps - ADI_GOLD[0]) < ADI_GOLD[1], eps


def test_oracle_jac3d_f32_kat_golden():
    eps = H.oracle().orc_jac3d_f32_kat(384, 100, 0, None)
    assert abs(eps - JAC3_GOLD[0]) < JAC3_GOLD[1], eps


def test_oracle_jac2d_f32_kat_golden():
    eps = H.oracle().orc_jac2d_f32_kat(8000, 100, 0, None)
    assert abs(eps - JAC2_GOLD[0]) < JAC2_GOLD[1], eps


# ---------------------------------------------------------------- host path through the ABI
def test_adikat_host_path_matches_oracle():
    n, its = 22, 6
    eps = H.run_adikat(n, its, maxeps=0.0, targets=D.DEVICE_TYPE_HOST)
    ref = np.zeros(its)
    H.oracle().orc_adi_kat(n, its, 0, H.vp(ref))
    assert np.array_equal(eps, ref)


@pytest.mark.parametrize('rank,n', [(2, 40), (3, 14)])
def test_jackat_host_path_matches_oracle(rank, n):
    its = 5
    last, eps = H.run_jackat(rank, n, its, maxeps=0.0, targets=D.DEVICE_TYPE_HOST)
    ref = np.zeros(its, dtype=np.float32)
    f = H.oracle().orc_jac2d_f32_kat if rank == 2 else H.oracle().orc_jac3d_f32_kat
    f(n, its, 0, H.vp(ref))
    assert np.array_equal(eps, ref)


# ---------------------------------------------------------------- GPU end to end
def _gpu(fn, *a, **kw):
    before = D.stats()
    r = fn(*a, **kw)
    after = D.stats()
    assert after['loops_gpu'] > before['loops_gpu'] and after['loops_host'] == before['loops_host']
    return r


@pytest.mark.gpu
def test_adikat_gpu_golden():
    eps = _gpu(H.run_adikat, 384, 100)
    assert len(eps) == 100
    assert abs(eps[-1] - ADI_GOLD[0]) < ADI_GOLD[1], eps[-1]
    ref = np.zeros(100)
    H.oracle().orc_adi_kat(384, 100, 0, H.vp(ref))
    assert np.array_equal(eps, ref)


@pytest.mark.gpu
def test_jac3kat_gpu_golden():
    last, eps = _gpu(H.run_jackat, 3, 384, 100)
    assert abs(last - JAC3_GOLD[0]) < JAC3_GOLD[1], last
    ref = np.zeros(100, dtype=np.float32)
    H.oracle().orc_jac3d_f32_kat(384, 100, 0, H.vp(ref))
    assert np.array_equal(eps, ref)


@pytest.mark.gpu
def test_jac2kat_gpu_golden():
    last, eps = _gpu(H.run_jackat, 2, 8000, 100)
    assert abs(last - JAC2_GOLD[0]) < JAC2_GOLD[1], last
    ref = np.zeros(100, dtype=np.float32)
    H.oracle().orc_jac2d_f32_kat(8000, 100, 0, H.vp(ref))
    assert np.array_equal(eps, ref)
