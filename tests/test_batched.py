"""Batched uneven grids: many independent 2-D Jacobi grids, each its own DVM
arrays, regions and loops.  The per-grid form runs every grid's loops through
their ordinary handlers; the batched form queues one iteration's loops between
dvmh_amd_batch_begin_C / dvmh_amd_batch_end_C and runs each loop site as one
persistent kernel over all grids' tiles -- or, with fused loop pairs on (the
default), runs each grid's stencil loop and copy + MAX loop as one pair, all
grids' pairs through one call of the pair handler.  Every form is compared bit
for bit with the oracle (grid values) and exactly (per-grid MAX eps)."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H


def _check(res, ref_list, iters):
    for ref in ref_list:
        g = ref['g']
        got = res['grids'][g]
        assert H.bits_equal(got['A'], ref['A']), f'grid {g} (side {res["sides"][g]}): A differs'
        assert H.bits_equal(got['B'], ref['B']), f'grid {g}: B differs'
        assert np.array_equal(res['eps'][:, g], ref['eps']), f'grid {g}: eps differs'


def test_sides_follow_seed():
    """The program draws the same sides as the oracle, all within [64, 1024]."""
    s = H.batched_sides(512)
    assert s.min() >= 64 and s.max() <= 1024
    res = H.run_batched(6, 0, 0, targets=D.DEVICE_TYPE_HOST, fetch=False)
    assert np.array_equal(res['sides'], s[:6])


@pytest.mark.parametrize('batched', [0, 1])
def test_host_path_bit_exact(batched):
    """Host handlers (batching only applies to GPU loops, so batched=1 must be a no-op here)."""
    sides = [5, 2, 3, 17, 64]
    res = H.run_batched(len(sides), 4, batched, sides=sides, targets=D.DEVICE_TYPE_HOST)
    _check(res, H.oracle_batched(sides, 4), 4)


def _pair_stats():
    out = (ctypes.c_long * 2)()
    D.runtime().dvmh_amd_pair_stats_C(out)
    return list(out)


def _gpu_run(count, iters, batched, **kw):
    before = D.stats()
    pb = _pair_stats()
    res = H.run_batched(count, iters, batched, **kw)
    after = D.stats()
    pa = _pair_stats()
    assert after['gpu_present'] == 1
    assert after['loops_host'] == before['loops_host']
    d = {k: after[k] - before[k] for k in after}
    d['pair_calls'], d['pairs'] = pa[0] - pb[0], pa[1] - pb[1]
    return res, d


class _Fuse:
    """Fused loop pairs on or off for a block."""

    def __init__(self, on):
        self.on = on
        self.f = D.runtime().dvmh_amd_set_fuse_C
        self.f.argtypes = [ctypes.c_long]

    def __enter__(self):
        self.f(self.on)

    def __exit__(self, *exc):
        self.f(1)


@pytest.mark.gpu
@pytest.mark.parametrize('fuse', [0, 1])
@pytest.mark.parametrize('batched', [0, 1])
def test_gpu_ragged_bit_exact(batched, fuse):
    """Ragged sides including a grid with an empty interior (side 2), a single
    interior point (side 3) and sides that are not multiples of any tile."""
    sides = [2, 3, 67, 5, 130, 64, 1, 257, 33, 1024]
    sides = [max(s, 2) for s in sides]
    iters = 5
    with _Fuse(fuse):
        res, d = _gpu_run(len(sides), iters, batched, sides=sides)
    _check(res, H.oracle_batched(sides, iters), iters)
    nonempty = sum(1 for s in sides if s > 2)
    if batched and not fuse:
        # two fused launches per iteration (stencil site, copy+MAX site)
        assert d['batch_groups'] == 2 * iters
        assert d['batch_loops'] == 2 * iters * nonempty
    elif batched:
        # one pair-handler call per iteration for all grids' pairs
        assert d['batch_groups'] == 0 and d['pair_calls'] == iters and d['pairs'] == iters * nonempty
    else:
        assert d['batch_groups'] == 0 and d['pair_calls'] == 0


KIND_BATCH = 4
# (block, tune): the per-loop default shape, a one-row shape with eight rows per
# thread, a narrow shape with two rows per thread (direct tiles), then LDS march
# tiles of the stencil loop (VARIANT 1, TILE_A points per thread, CHUNK rows per
# tile; the copy + MAX loop ignores the variant)
BATCH_FORCED = [((32, 8, 1), (1, 0, 0, 0, 0, 0, 0, 0)), ((256, 1, 1), (8, 1, 0, 0, 0, 0, 0, 0)),
                ((64, 2, 1), (2, 0, 0, 0, 0, 0, 0, 0)), ((128, 1, 1), (1, 0, 1, 16, 2, 0, 0, 0)),
                ((256, 1, 1), (1, 0, 1, 8, 1, 0, 0, 0)), ((64, 1, 1), (1, 0, 1, 32, 4, 0, 0, 0)),
                ((512, 1, 1), (1, 0, 1, 64, 1, 0, 0, 0))]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', BATCH_FORCED)
def test_gpu_batch_configs_bit_exact(cfg):
    """Batch kernels under forced configurations other than the batch
    default: the same ragged grids, bit-exact against the oracle."""
    import ctypes
    (block, tune) = cfg
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    b = (ctypes.c_long * 3)(*block)
    t = (ctypes.c_long * 8)(*tune)
    r.dvmh_amd_tuner_force_C(KIND_BATCH, ctypes.cast(b, ctypes.c_void_p), ctypes.cast(t, ctypes.c_void_p))
    try:
        sides = [2, 3, 67, 5, 130, 64, 257, 33, 1024]
        with _Fuse(0):   # the batch kernels, not the pair handler
            res, d = _gpu_run(len(sides), 4, 1, sides=sides)
        _check(res, H.oracle_batched(sides, 4), 4)
        assert d['batch_groups'] == 8
    finally:
        r.dvmh_amd_tuner_force_C(KIND_BATCH, None, None)


@pytest.mark.gpu
@pytest.mark.parametrize('fuse', [0, 1])
def test_gpu_batched_matches_per_grid_seeded(fuse):
    """64 seeded grids, 6 iterations: batched and per-grid forms agree bit for
    bit with each other and with the oracle on a sample of grids."""
    count, iters = 64, 6
    a, _ = _gpu_run(count, iters, 0)
    with _Fuse(fuse):
        b, d = _gpu_run(count, iters, 1)
    if fuse:
        assert d['pair_calls'] == iters and d['pairs'] == iters * count
    else:
        assert d['batch_groups'] == 2 * iters and d['batch_loops'] == 2 * iters * count
    assert np.array_equal(a['eps'], b['eps'])
    for g in range(count):
        assert H.bits_equal(a['grids'][g]['A'], b['grids'][g]['A'])
    sample = [0, 1, 17, 63, int(np.argmax(a['sides'])), int(np.argmin(a['sides']))]
    _check(b, H.oracle_batched(a['sides'], iters, grids=sample), iters)


class _Graphs:
    """Graph replay of repeating batch queues on (DVMH_AMD_GRAPH) for a block."""

    def __init__(self):
        self.f = D.runtime().dvmh_amd_set_graph_C
        self.f.argtypes = [ctypes.c_long]

    def __enter__(self):
        self.f(1)

    def __exit__(self, *exc):
        self.f(0)


@pytest.mark.gpu
@pytest.mark.parametrize('fuse', [0, 1])
def test_gpu_batched_graph_replay_bit_exact(fuse):
    """The batched time step captured as one graph: the first iteration runs
    eagerly, the second is captured (and launched), the rest replay the graph.
    Results equal the eager batched form and the oracle bit for bit."""
    count, iters = 64, 8
    a, _ = _gpu_run(count, iters, 1)
    with _Graphs(), _Fuse(fuse):
        b, d = _gpu_run(count, iters, 1)
    assert d['graph_replays'] == iters - 2
    # only the eager and the captured iteration run the handlers
    if fuse:
        assert d['pair_calls'] == 2 and d['batch_groups'] == 0
    else:
        assert d['batch_groups'] == 2 * 2
    assert np.array_equal(a['eps'], b['eps'])
    for g in range(count):
        assert H.bits_equal(a['grids'][g]['A'], b['grids'][g]['A'])
        assert H.bits_equal(a['grids'][g]['B'], b['grids'][g]['B'])
    sample = [0, 5, 63, int(np.argmax(a['sides'])), int(np.argmin(a['sides']))]
    _check(b, H.oracle_batched(a['sides'], iters, grids=sample), iters)


@pytest.mark.gpu
def test_gpu_batched_graph_ragged():
    """Ragged sides (empty and one-point interiors) with graph replay, twice in
    a row on fresh arrays (the graphs of the first run are dropped with its
    arrays)."""
    sides = [2, 3, 67, 5, 130, 64, 2, 257, 33, 1024]
    iters = 6
    with _Graphs():
        for _ in range(2):
            res, d = _gpu_run(len(sides), iters, 1, sides=sides)
            assert d['graph_replays'] == iters - 2
            _check(res, H.oracle_batched(sides, iters), iters)


@pytest.mark.gpu
@pytest.mark.parametrize('graph', [0, 1])
def test_gpu_full_workload_properties(graph):
    """512 seeded grids x 200 iterations (the workload's size), batched (and
    with the time step replayed as a graph): every grid's eps is
    non-increasing, the boundaries stay 1.0, and the smallest and largest
    grids match the oracle bit for bit."""
    count, iters = 512, 200
    if graph:
        with _Graphs():
            res, d = _gpu_run(count, iters, 1, fetch=True)
        assert d['graph_replays'] == iters - 2
    else:
        res, d = _gpu_run(count, iters, 1, fetch=True)
        assert d['pairs'] == iters * count
    assert np.all(np.diff(res['eps'], axis=0) <= 0)
    for g in range(count):
        n = int(res['sides'][g])
        A = res['grids'][g]['A'].reshape(n, n)
        assert np.all(A[0] == 1.0) and np.all(A[-1] == 1.0) and np.all(A[:, 0] == 1.0) and np.all(A[:, -1] == 1.0)
    sample = [int(np.argmin(res['sides'])), int(np.argmax(res['sides']))]
    _check(res, H.oracle_batched(res['sides'], iters, grids=sample), iters)
