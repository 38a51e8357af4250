"""Chains of ACROSS sweeps: consecutive executions of the SOR sweep loop
queued in one batch (programs/sor3d.c sor3d_iterate_chain) run through the
loop's chain handler, every sweep two tile-planes behind the one before it in
one launch sequence, each with its own REDUCTION(MAX) segment.  Results must
equal the sequential sweeps bit for bit, whatever the group size, tile shape
and grid size; configurations other than the tiled wavefront fall back to one
sweep after another."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

KIND_WAVEFRONT = 2
KIND_CHAIN = 5


def _force(block, tune):
    """The same configuration for the loop's own sweeps and for its chains
    (the chain handler asks its own tuner kind)."""
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    for kind in (KIND_WAVEFRONT, KIND_CHAIN):
        if block is None:
            r.dvmh_amd_tuner_force_C(kind, None, None)
            continue
        b = np.array(block, dtype=np.int64)
        t = np.array(tune, dtype=np.int64)
        r.dvmh_amd_tuner_force_C(kind, H.vp(b), H.vp(t))


def _tile(nt, ta, tc):
    tune = [0] * 8
    tune[0] = 1      # DEPTH
    tune[2] = 1      # VARIANT: tiled wavefront
    tune[4] = ta     # TILE_A
    tune[5] = tc     # TILE_B
    return (nt, 1, 1), tune


def test_chain_host_path():
    """Without a GPU the regions run on the host: the batch queues nothing and
    every sweep runs in order."""
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    res = H.run_sor3d_chain(12, 5, 3)
    ref = H.oracle_sor3d(12, 5)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', [(64, 4, 8), (256, 8, 8), (256, 8, 16), (512, 16, 16), (256, 12, 16), (128, 16, 16), (32, 16, 8)])
@pytest.mark.parametrize('n,iters,group', [(5, 3, 3), (23, 7, 3), (23, 6, 6), (41, 9, 4), (70, 5, 5)])
def test_chain_bit_exact(cfg, n, iters, group):
    """Tiles overhanging the loop, tiles larger than it (n = 5), groups that
    divide the sweeps and groups that leave a shorter last group; tiles with at
    most one column per thread, up to four (128 x 16 x 16) and more (32 x 16 x 8)."""
    block, tune = _tile(*cfg)
    _force(block, tune)
    before = D.stats()
    try:
        res = H.run_sor3d_chain(n, iters, group)
    finally:
        _force(None, None)
    after = D.stats()
    ref = H.oracle_sor3d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['eps'], ref['eps'])
    assert after['batch_groups'] > before['batch_groups']


@pytest.mark.gpu
def test_chain_falls_back_without_tiles():
    """Hyperplane configuration (not the tiled wavefront): the chain handler
    runs the group's sweeps one after another."""
    _force((64, 1, 1), [1, 0, 0, 0, 0, 0, 0, 0])
    try:
        res = H.run_sor3d_chain(19, 5, 5)
    finally:
        _force(None, None)
    ref = H.oracle_sor3d(19, 5)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', [(256, 8, 8), (256, 12, 16)])
def test_chain_full_size(cfg):
    """320^3 (the SOR workload's size), 6 sweeps in one chain: 8^3 tiles, and
    12 x 12 x 16 tiles on 256 threads (one column per thread, the shape the
    tuner keeps)."""
    block, tune = _tile(*cfg)
    _force(block, tune)
    try:
        res = H.run_sor3d_chain(320, 6, 6)
    finally:
        _force(None, None)
    ref = H.oracle_sor3d(320, 6)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
def test_chain_tuned_bit_exact():
    """Cost-model mode: the per-loop site is explored sweep by sweep inside the
    first groups, then the chain's own site times its tile shapes group by
    group (exhaustively: no records of chains train the model).  Every run is
    bit-exact and the chain site ends settled with a recorded best shape."""
    from dvmh_amd import costmodel as CM
    r = CM.rt()
    if r.dvmh_amd_model_load_C(CM.MODEL.encode()) < 0:
        pytest.skip('no trained cost model')
    r.dvmh_amd_set_sync_mode_C.argtypes = [ctypes.c_long]
    r.dvmh_amd_tuner_query_chain_C.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    r.dvmh_amd_tuner_query_chain_C.restype = ctypes.c_long
    n, iters, group = 40, 10, 5
    ref = H.oracle_sor3d(n, iters)
    r.dvmh_amd_tuner_clear_C()
    r.dvmh_amd_tuner_set_mode_C(2)
    r.dvmh_amd_set_sync_mode_C(1)
    runs = 0
    try:
        for runs in range(1, 80):
            res = H.run_sor3d_chain(n, iters, group)
            assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps']), runs
            if r.dvmh_amd_tuner_site_pending_C(b'sor3d.cdv', 32) == 0:
                break
        pending = r.dvmh_amd_tuner_site_pending_C(b'sor3d.cdv', 32)
        q = (ctypes.c_long * 13)()
        found = r.dvmh_amd_tuner_query_chain_C(b'sor3d.cdv', 32, q)
    finally:
        r.dvmh_amd_tuner_set_mode_C(0)
        r.dvmh_amd_set_sync_mode_C(0)
        r.dvmh_amd_tuner_clear_C()
    assert pending == 0, (pending, runs)
    assert found == 1 and q[5 + 2] == 1   # a tiled-wavefront shape kept for the chain
