"""The 64-bit index kernels (dvmh_loop_guess_index_type_C answering rt_LONG,
rts/dvmh/dvmh2.cpp:4097-4154).  No tested array has 2^31 elements, so the
runtime is told to hand every loop 64-bit indices (dvmh_amd_set_index64_C)
and each program runs its long long kernels -- PARALLEL, LDS march, line and
register sweeps, hyperplanes, tiled wavefronts and batch kernels -- on grids
small enough to check against the oracle bit for bit."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H


def _rt():
    r = D.runtime()
    r.dvmh_amd_set_index64_C.argtypes = [ctypes.c_long]
    r.dvmh_loop_guess_index_type_C.restype = ctypes.c_long
    return r


class _Index64:
    def __enter__(self):
        _rt().dvmh_amd_set_index64_C(1)
        self.before = D.stats()['loops_gpu']
        return self

    def __exit__(self, *exc):
        _rt().dvmh_amd_set_index64_C(0)
        self.gpu_loops = D.stats()['loops_gpu'] - self.before
        return False


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(3, 2), (37, 3), (130, 2)])
def test_index64_jacobi(n, iters):
    with _Index64() as f:
        r2 = H.run_jac2d(n, iters)
        r3 = H.run_jac3d(n if n < 100 else 66, iters)
    ref2 = H.oracle_jac2d(n, iters)
    ref3 = H.oracle_jac3d(n if n < 100 else 66, iters)
    assert f.gpu_loops > 0
    assert H.bits_equal(r2['A'], ref2['A']) and H.bits_equal(r2['B'], ref2['B']) and np.array_equal(r2['eps'], ref2['eps'])
    assert H.bits_equal(r3['A'], ref3['A']) and np.array_equal(r3['eps'], ref3['eps']) and H.sum_close(r3['sum'], ref3['sum'])


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(5, 2), (23, 3), (41, 2)])
def test_index64_across(n, iters):
    """SOR (three dependent axes: hyperplanes or tiled wavefront) and ADI
    (one dependent axis per sweep: line kernels) with 64-bit indices."""
    with _Index64() as f:
        rs = H.run_sor3d(n, iters)
        ra = H.run_adi3d(n, iters)
    refs = H.oracle_sor3d(n, iters)
    refa = H.oracle_adi3d(n, iters)
    assert f.gpu_loops > 0
    assert H.bits_equal(rs['A'], refs['A']) and np.array_equal(rs['eps'], refs['eps'])
    assert H.bits_equal(ra['A'], refa['A']) and np.array_equal(ra['max'], refa['max']) and np.array_equal(ra['loc'], refa['loc'])


@pytest.mark.gpu
def test_index64_batched():
    """Batch kernels of uneven grids with 64-bit indices (the batch handler asks
    each instance's index type)."""
    sides = [2, 3, 17, 64, 130]
    with _Index64() as f:
        res = H.run_batched(len(sides), 3, True, sides=sides)
    assert f.gpu_loops > 0
    for r in H.oracle_batched(np.array(sides), 3):
        g = r['g']
        assert H.bits_equal(res['grids'][g]['A'], r['A']) and H.bits_equal(res['grids'][g]['B'], r['B'])
        assert np.array_equal(res['eps'][:, g], r['eps'])


@pytest.mark.gpu
@pytest.mark.parametrize('kind,block,tune', [
    (2, (256, 1, 1), [1, 0, 1, 0, 8, 8, 0, 0]),     # tiled wavefront, 8 x 8 x 8 tiles
    (2, (64, 1, 1), [1, 0, 0, 0, 0, 0, 0, 0]),      # one launch per hyperplane
    (1, (64, 1, 1), [1, 0, 0, 0, 0, 0, 0, 0]),      # direct line kernel
])
def test_index64_forced_variants(kind, block, tune):
    """Kernel variants that keep a 64-bit instantiation, forced by kind."""
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    b = np.array(block, dtype=np.int64)
    t = np.array(tune, dtype=np.int64)
    r.dvmh_amd_tuner_force_C(kind, H.vp(b), H.vp(t))
    try:
        with _Index64():
            rs = H.run_sor3d(19, 2)
            ra = H.run_adi3d(19, 2)
    finally:
        r.dvmh_amd_tuner_force_C(kind, None, None)
    refs = H.oracle_sor3d(19, 2)
    refa = H.oracle_adi3d(19, 2)
    assert H.bits_equal(rs['A'], refs['A']) and np.array_equal(rs['eps'], refs['eps'])
    assert H.bits_equal(ra['A'], refa['A']) and np.array_equal(ra['max'], refa['max']) and np.array_equal(ra['loc'], refa['loc'])
