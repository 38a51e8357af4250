"""Execution forms that change when results are synchronised and how launches
are issued, never the arithmetic: regions replayed as captured HIP graphs
(REGION_AMD_GRAPH + dvmh_amd_set_graph_C(1)), results synchronised only when
needed (sync mode 1), and the tuner's exhaustive and cost-model modes.  Every
form must match the oracle bit for bit (SUM within 1e-12)."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H
from dvmh_amd import costmodel as CM


def _rt():
    r = CM.rt()
    r.dvmh_amd_set_graph_C.argtypes = [ctypes.c_long]
    r.dvmh_amd_set_sync_mode_C.argtypes = [ctypes.c_long]
    r.dvmh_amd_set_graph_min_nodes_C.argtypes = [ctypes.c_long]
    return r


class _Form:
    def __init__(self, graph=0, sync=0, mode=0):
        self.graph, self.sync, self.mode = graph, sync, mode

    def __enter__(self):
        r = _rt()
        r.dvmh_amd_set_graph_C(self.graph)
        r.dvmh_amd_set_graph_min_nodes_C(0)   # replay every captured region, however small
        r.dvmh_amd_set_sync_mode_C(self.sync)
        r.dvmh_amd_tuner_set_mode_C(self.mode)
        r.dvmh_amd_tuner_clear_C()
        self.before = D.stats()
        return self

    def __exit__(self, *exc):
        r = _rt()
        self.after = D.stats()
        r.dvmh_amd_set_graph_C(0)
        r.dvmh_amd_set_graph_min_nodes_C(-1)
        r.dvmh_amd_set_sync_mode_C(0)
        r.dvmh_amd_tuner_set_mode_C(0)
        r.dvmh_amd_tuner_clear_C()
        return False

    def replays(self):
        return self.after['graph_replays'] - self.before['graph_replays']


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(37, 9), (258, 12)])
def test_jac2d_graph_replay(n, iters):
    with _Form(graph=1, sync=1) as f:
        res = H.run_jac2d(n, iters, flags=D.REGION_AMD_GRAPH)
    ref = H.oracle_jac2d(n, iters)
    assert H.bits_equal(res['A'], ref['A']) and H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])
    # the first iterations run eagerly, the repeating region is captured, then replayed
    assert f.replays() >= iters - 4


@pytest.mark.gpu
def test_jac3d_graph_replay():
    n, iters = 66, 10
    with _Form(graph=1, sync=1) as f:
        res = H.run_jac3d(n, iters, flags=D.REGION_AMD_GRAPH)
    ref = H.oracle_jac3d(n, iters)
    assert H.bits_equal(res['A'], ref['A']) and H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])
    assert H.sum_close(res['sum'], ref['sum'])
    assert f.replays() > 0


@pytest.mark.gpu
def test_sor3d_graph_replay():
    n, iters = 34, 8
    with _Form(graph=1, sync=1) as f:
        res = H.run_sor3d(n, iters, flags=D.REGION_AMD_GRAPH)
    ref = H.oracle_sor3d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['eps'], ref['eps'])
    assert f.replays() > 0


@pytest.mark.gpu
def test_adi3d_graph_replay():
    n, steps = 40, 6
    with _Form(graph=1, sync=1) as f:
        res = H.run_adi3d(n, steps, flags=D.REGION_AMD_GRAPH)
    ref = H.oracle_adi3d(n, steps)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])
    assert f.replays() > 0


@pytest.mark.gpu
@pytest.mark.parametrize('mode', [1, 2])
def test_tuned_forms_bit_exact(mode):
    """Exhaustive sweep (every candidate configuration runs at least once) and
    the cost-model form: results identical to the oracle."""
    if mode == 2 and CM.rt().dvmh_amd_model_load_C(CM.MODEL.encode()) < 0:
        pytest.skip('no trained cost model')
    with _Form(sync=1, mode=mode):
        r3 = H.run_jac3d(34, 40)
        r2 = H.run_jac2d(70, 40)
        ra = H.run_adi3d(24, 30)
        rs = H.run_sor3d(20, 30)
    ref3 = H.oracle_jac3d(34, 40)
    assert H.bits_equal(r3['A'], ref3['A']) and np.array_equal(r3['eps'], ref3['eps']) and H.sum_close(r3['sum'], ref3['sum'])
    ref2 = H.oracle_jac2d(70, 40)
    assert H.bits_equal(r2['A'], ref2['A']) and np.array_equal(r2['eps'], ref2['eps'])
    refa = H.oracle_adi3d(24, 30)
    assert H.bits_equal(ra['A'], refa['A']) and np.array_equal(ra['max'], refa['max']) and np.array_equal(ra['loc'], refa['loc'])
    refs = H.oracle_sor3d(20, 30)
    assert H.bits_equal(rs['A'], refs['A']) and np.array_equal(rs['eps'], refs['eps'])


@pytest.mark.gpu
def test_graph_replay_interleaved_with_eager_work():
    """Replays interleaved with eager regions of another program and host
    synchronisations: every replay still matches (captured reductions are
    re-posted on every replay)."""
    n, iters = 50, 6
    with _Form(graph=1, sync=1) as f:
        a = H.run_jac2d(n, iters, flags=D.REGION_AMD_GRAPH)
        b = H.run_jac3d(20, 3)           # eager, between two graph runs
        c = H.run_jac2d(n, iters, flags=D.REGION_AMD_GRAPH)
    ref = H.oracle_jac2d(n, iters)
    for res in (a, c):
        assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])
    ref3 = H.oracle_jac3d(20, 3)
    assert H.bits_equal(b['A'], ref3['A'])
    assert f.replays() > 0


@pytest.mark.gpu
@pytest.mark.parametrize('prog,n', [('sor3d', 320), ('adi3d', 256)])
def test_cost_model_form_at_full_size(prog, n):
    """The cost-model form at the size the SOR (320^3) and ADI (256^3)
    workloads are timed.  A first run explores the model's trial set (every
    sweep may use another configuration) and must already match the oracle;
    once no site is pending, a fresh run replays its repeating region as a
    graph with the configurations the tuner kept, and matches as well."""
    if CM.rt().dvmh_amd_model_load_C(CM.MODEL.encode()) < 0:
        pytest.skip('no trained cost model')
    run = {'sor3d': H.run_sor3d, 'adi3d': H.run_adi3d}[prog]
    oracle = {'sor3d': H.oracle_sor3d, 'adi3d': H.oracle_adi3d}[prog]
    explore, final = 10, 4

    def same(res, ref):
        if prog == 'sor3d':
            return H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])
        return (H.bits_equal(res['A'], ref['A']) and np.array_equal(res['max'], ref['max'])
                and np.array_equal(res['loc'], ref['loc']))

    r = _rt()
    with _Form(graph=1, sync=1, mode=2) as f:
        first = run(n, explore, flags=D.REGION_AMD_GRAPH)
        for _ in range(6):
            if r.dvmh_amd_tuner_pending_C() == 0:
                break
            run(n, explore, flags=D.REGION_AMD_GRAPH)
        pending = r.dvmh_amd_tuner_pending_C()
        replays0 = D.stats()['graph_replays']
        second = run(n, final, flags=D.REGION_AMD_GRAPH)
        replays = D.stats()['graph_replays'] - replays0
    assert same(first, oracle(n, explore))
    assert pending == 0
    assert same(second, oracle(n, final))
    assert replays > 0


@pytest.mark.gpu
def test_small_regions_run_eagerly_by_default():
    """Under the default replay threshold a region whose captured graph has
    few nodes (2-D Jacobi's fused pair: three kernels) runs eagerly after its
    capture, while the ADI step (seven loops) is replayed; both bit-exact."""
    r = _rt()
    r.dvmh_amd_set_graph_C(1)
    r.dvmh_amd_set_sync_mode_C(1)
    try:
        s0 = D.stats()
        res = H.run_jac2d(130, 12, flags=D.REGION_AMD_GRAPH)
        s1 = D.stats()
        ra = H.run_adi3d(40, 6, flags=D.REGION_AMD_GRAPH)
        s2 = D.stats()
    finally:
        r.dvmh_amd_set_graph_C(0)
        r.dvmh_amd_set_sync_mode_C(0)
    ref = H.oracle_jac2d(130, 12)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])
    refa = H.oracle_adi3d(40, 6)
    assert H.bits_equal(ra['A'], refa['A']) and np.array_equal(ra['max'], refa['max'])
    assert s1['graph_replays'] == s0['graph_replays']
    assert s2['graph_replays'] > s1['graph_replays']
