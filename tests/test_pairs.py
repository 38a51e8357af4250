"""Fused loop pairs (programs/pairs.cdv, converted by the front end): a stencil
loop and the loop after it that touches every array only at the loop point run
as one pass through the pair handler the emitter generates (emit/hipgen.py
pair_code), the second loop at tile-edge points in a second kernel.

The second loop is not idempotent (A = A * 0.5 + B, C = C + 1, cnt = cnt + 1),
so a point skipped or run twice changes A, C and cnt.  Results are compared bit
for bit with a numpy restatement of the same operation order, with the
loop-by-loop GPU run (batched = 0, no pair handler) and, on the CPU, with the
host path.  C must equal the number of steps at every interior point and 0 on
the boundary, cnt the number of interior points; MAXLOC ties keep the lower
loc as on the host (strict > in the body, increasing loc)."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
from test_cdvmh import _u01

_dp = ctypes.POINTER(ctypes.c_double)
KIND_PAIR = 6


def reference(shape, steps, seed):
    """numpy restatement of pairs2 / pairs3: per step eps, loc, cnt, bmin, then A, C."""
    n = int(np.prod(shape))
    A = _u01(seed, np.arange(n, dtype=np.uint64)).reshape(shape)
    B = np.zeros(shape)
    C = np.zeros(shape)
    inner = tuple(slice(1, s - 1) for s in shape)
    hist = []
    with np.errstate(all='ignore'):
        for _ in range(steps):
            if len(shape) == 2:
                s = A[:-2, 1:-1] + A[2:, 1:-1] + A[1:-1, :-2] + A[1:-1, 2:]
                B[inner] = s * 0.25 - A[inner] * 0.125
            else:
                s = A[:-2, 1:-1, 1:-1] + A[2:, 1:-1, 1:-1] + A[1:-1, :-2, 1:-1] + A[1:-1, 2:, 1:-1] + A[1:-1, 1:-1, :-2] + A[1:-1, 1:-1, 2:]
                B[inner] = s / 6.0 - A[inner] * 0.125
            npts = int(np.prod([max(0, x - 2) for x in shape]))
            if npts:
                d = np.abs(B[inner] - A[inner])
                eps = d.max()
                idx = np.array(np.unravel_index(np.argmax(d), d.shape)) + 1   # first maximum: lowest loc
                loc = float(np.ravel_multi_index(tuple(idx), shape))
                bmin = B[inner].min()
            else:
                eps, loc, bmin = -1.0, -1.0, 1.0e300
            A[inner] = A[inner] * 0.5 + B[inner]
            C[inner] = C[inner] + 1.0
            hist += [eps, loc, float(npts), bmin]
    return np.array(hist), A.ravel(), C.ravel()


def run(shape, steps, batched, seed):
    lib = D.load('libcdv_pairs.so')
    n = int(np.prod(shape))
    hist = np.full(4 * steps, np.nan)
    a = np.zeros(n)
    c = np.zeros(n)
    if len(shape) == 2:
        f = lib.pairs2
        f.argtypes = [ctypes.c_long] * 3 + [ctypes.c_int, ctypes.c_ulonglong, _dp, _dp, _dp]
        f(shape[0], shape[1], steps, batched, seed, hist.ctypes.data_as(_dp), a.ctypes.data_as(_dp), c.ctypes.data_as(_dp))
    else:
        f = lib.pairs3
        f.argtypes = [ctypes.c_long] * 4 + [ctypes.c_int, ctypes.c_ulonglong, _dp, _dp, _dp]
        f(shape[0], shape[1], shape[2], steps, batched, seed, hist.ctypes.data_as(_dp), a.ctypes.data_as(_dp),
          c.ctypes.data_as(_dp))
    return hist, a, c


def pair_stats():
    r = D.runtime()
    out = (ctypes.c_long * 2)()
    r.dvmh_amd_pair_stats_C(out)
    return list(out)


def check(shape, steps, batched, seed=D.SEED_BASE + 60):
    hist, a, c = run(shape, steps, batched, seed)
    rh, ra, rc = reference(shape, steps, seed)
    assert np.array_equal(c, rc), 'a point ran the second loop other than once per step'
    assert np.array_equal(hist.view(np.uint64), rh.view(np.uint64))
    assert np.array_equal(a.view(np.uint64), ra.view(np.uint64))


def test_reference_counts_and_residual():
    """The restatement: C counts steps on the interior only, cnt the interior
    points; eps is where |B - A| peaks and loc the first such point."""
    h, a, c = reference((5, 7), 3, 11)
    cc = c.reshape(5, 7)
    assert (cc[1:-1, 1:-1] == 3).all() and cc[0].sum() == 0 and cc[:, -1].sum() == 0
    assert h[2] == 15.0
    h3, _, c3 = reference((4, 5, 6), 2, 12)
    assert h3[2] == 2 * 3 * 4 and c3.sum() == 2 * 24


def test_pairs_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    for shape in ((5, 9), (3, 3), (6, 5, 7)):
        check(shape, 3, 1)
        check(shape, 3, 0)


def test_pair_handlers_registered():
    """The converted library registers one pair handler per fusable pair and
    exports it (the runtime calls it through the registration)."""
    lib = D.load('libcdv_pairs.so')
    for name in ('loop_pairs_cdv_53__loop_pairs_cdv_60_cuda_pair', 'loop_pairs_cdv_130__loop_pairs_cdv_138_cuda_pair'):
        assert hasattr(lib, name)


# ---- GPU

def _force(block, tune):
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    if block is None:
        r.dvmh_amd_tuner_force_C(KIND_PAIR, None, None)
        return
    b = (ctypes.c_long * 3)(*block)
    t = (ctypes.c_long * 8)(*tune)
    r.dvmh_amd_tuner_force_C(KIND_PAIR, ctypes.cast(b, ctypes.c_void_p), ctypes.cast(t, ctypes.c_void_p))


# (block, march, points per thread): every tile edge kind, a march of two
# planes (every plane on an edge), wide and narrow tiles
SHAPES2 = [((64, 1, 1), 2, 1), ((128, 1, 1), 8, 2), ((256, 1, 1), 64, 4), ((1024, 1, 1), 16, 1), ((64, 1, 1), 5, 3)]
SHAPES3 = [((32, 2, 1), 2, 1), ((64, 4, 1), 8, 2), ((32, 8, 1), 16, 4), ((128, 2, 1), 3, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize('shape', [(3, 3), (5, 9), (37, 300), (130, 1029), (301, 67)])
@pytest.mark.parametrize('cfg', SHAPES2)
def test_pairs2_fused(shape, cfg):
    block, march, cols = cfg
    _force(block, [1, 0, 1, march, cols, 0, 0, 0])
    try:
        before = pair_stats()
        check(shape, 3, 1)
        after = pair_stats()
        assert after[1] - before[1] == 3, 'the pairs did not run through the pair handler'
        check(shape, 3, 0)
        assert pair_stats() == after
    finally:
        _force(None, None)


@pytest.mark.gpu
@pytest.mark.parametrize('shape', [(3, 3, 3), (6, 5, 7), (20, 33, 70), (41, 18, 129)])
@pytest.mark.parametrize('cfg', SHAPES3)
def test_pairs3_fused(shape, cfg):
    block, march, cols = cfg
    _force(block, [1, 1, 1, march, cols, 0, 0, 0])
    try:
        before = pair_stats()
        check(shape, 2, 1)
        assert pair_stats()[1] - before[1] == 2
    finally:
        _force(None, None)


@pytest.mark.gpu
def test_pairs_default_and_tuned():
    """Default pair configuration, then the exhaustive tuner over the pair
    site's candidates: every explored configuration stays bit-exact."""
    r = D.runtime()
    check((200, 333), 4, 1)
    r.dvmh_amd_tuner_set_mode_C(1)
    try:
        for _ in range(3):
            check((64, 130), 6, 1)
    finally:
        r.dvmh_amd_tuner_set_mode_C(0)


@pytest.mark.gpu
def test_fuse_off_runs_loop_by_loop():
    r = D.runtime()
    r.dvmh_amd_set_fuse_C(0)
    try:
        before = pair_stats()
        check((40, 50), 2, 1)
        assert pair_stats() == before
    finally:
        r.dvmh_amd_set_fuse_C(1)


# ---- the emitter's pair analysis (CPU)

def _loops(*descs):
    import os
    import sys
    sys.path.insert(0, os.path.join(D.PKG_DIR, 'emit'))
    import hipgen
    return hipgen, [hipgen.Loop(d, 't') for d in descs]


def _stencil(body='B[i][j] = 0.25 * (A[i - 1][j] + A[i + 1][j] + A[i][j - 1] + A[i][j + 1]);', params=None, **kw):
    d = {'name': 'l1', 'vars': [('i', 'int'), ('j', 'int')], 'stencil': ['A'],
         'params': params or [('array', 'A', 'double', 2), ('array', 'B', 'double', 2)], 'body': body}
    d.update(kw)
    return d


def _pointwise(body='double d = fabs(B[i][j] - A[i][j]);\neps = Max(d, eps);\nA[i][j] = B[i][j];', params=None, **kw):
    d = {'name': 'l2', 'vars': [('i', 'int'), ('j', 'int')], 'reductions': [('max', 'eps', 'double')],
         'params': params or [('array', 'A', 'double', 2), ('array', 'B', 'double', 2)], 'body': body}
    d.update(kw)
    return d


def test_pair_eligible_accepts_time_step():
    hg, (a, b) = _loops(_stencil(), _pointwise())
    assert hg.pair_eligible(a, b)
    # loop variables of other names, an extra array and scalar in the second loop
    hg, (a, b) = _loops(_stencil(), _pointwise(body='C[p][q] = C[p][q] + B[p][q] * w;\nA[p][q] = B[p][q];',
                                              vars=[('p', 'int'), ('q', 'int')], reductions=[],
                                              params=[('array', 'A', 'double', 2), ('array', 'B', 'double', 2),
                                                      ('array', 'C', 'double', 2), ('scalar', 'w', 'double')]))
    assert hg.pair_eligible(a, b)


@pytest.mark.parametrize('what,l1,l2', [
    ('second loop reads a neighbour', {}, {'body': 'A[i][j] = B[i - 1][j];'}),
    ('second loop passes an array', {}, {'body': 'A[i][j] = g(B, i, j);'}),
    ('nested subscript', {}, {'body': 'A[i][j] = B[k[i]][j];'}),
    ('first loop reads a non-stencil array off the point', {'body': 'B[i][j] = A[i - 1][j] + B[i + 1][j];'}, {}),
    ('a name is an array in one loop and a scalar in the other', {},
     {'params': [('array', 'A', 'double', 2), ('scalar', 'B', 'double')], 'body': 'A[i][j] = B;'}),
    ('reduction variable named like a first-loop array', {}, {'reductions': [('max', 'B', 'double')],
                                                            'body': 'B = Max(B, A[i][j]);',
                                                            'params': [('array', 'A', 'double', 2)]}),
    ('across loop', {}, {'across': {'array': 'A', 'widths': [(1, 0), (0, 0)]}}),
    ('first loop without a stencil array', {'stencil': []}, {}),
])
def test_pair_eligible_rejects(what, l1, l2):
    hg, (a, b) = _loops(_stencil(**l1), _pointwise(**l2))
    assert not hg.pair_eligible(a, b), what


def test_pairs_registered_in_emitted_code(tmp_path):
    """emit() writes a pair handler and one registration per eligible
    consecutive pair, none for a pair that is not."""
    import os
    hg, _ = _loops(_stencil(), _pointwise())
    spec = {'program': 't', 'include': 'dvmh_prog_common.h',
            'loops': [_stencil(), _pointwise(), _pointwise(name='l3', body='A[i][j] = B[i + 1][j];', reductions=[])]}
    path = os.path.join(str(tmp_path), 't.loops')
    with open(path, 'w') as f:
        f.write(repr(spec))
    hg.emit(path, str(tmp_path))
    text = open(os.path.join(str(tmp_path), 't_hip.hip')).read()
    assert text.count('dvmh_amd_register_pair_C(') == 1
    assert 'l1__l2_cuda_pair' in text and 'l2__l3' not in text


def test_lds_reads_only_before_writes():
    """The fused second loop takes an array's value at its point from the first
    loop's LDS window only when every read of it comes before its first write
    in the body text (and the body has no loop); otherwise it reads memory."""
    hg, (a, b) = _loops(_stencil(), _pointwise())
    assert list(hg.lds_reads(a, b)) == ['A']
    hg, (a, b) = _loops(_stencil(), _pointwise(body='A[i][j] = A[i][j] * 0.5 + B[i][j];', reductions=[]))
    assert hg.lds_reads(a, b) == {}
    hg, (a, b) = _loops(_stencil(), _pointwise(body='double t = A[i][j];\nfor (int r = 0; r < 2; r++) t = t * 0.5;\nA[i][j] = t;',
                                              reductions=[]))
    assert hg.lds_reads(a, b) == {}
    hg, (a, b) = _loops(_stencil(), _pointwise(body='B[i][j] = A[i][j];', reductions=[]))
    assert list(hg.lds_reads(a, b)) == ['A']


@pytest.mark.gpu
def test_pairs_declined_run_loop_by_loop():
    """With 64-bit indices forced the pair handler declines (its kernels use
    32-bit indices): the runtime then runs each queued pair's two loops through
    their own handlers, in program order, with the same results."""
    r = D.runtime()
    r.dvmh_amd_set_index64_C.argtypes = [ctypes.c_long]
    r.dvmh_amd_set_index64_C(1)
    try:
        before = pair_stats()
        check((37, 300), 3, 1)
        check((6, 5, 7), 2, 1)
        assert pair_stats() == before
    finally:
        r.dvmh_amd_set_index64_C(0)


@pytest.mark.parametrize('body,written', [
    ('B[i][j] = c * A[i][j];', False),
    ('B[i][j] = (c == 2.0) ? A[i][j] : 0.0;', False),
    ('cc = 1.0; B[i][j] = cc * c;', False),
    ('p.c = 1.0; B[i][j] = c;', False),
    ('c = c + 1.0;', True),
    ('c += 2.0;', True),
    ('++c;', True),
    ('c--;', True),
    ('f(&c);', True),
])
def test_kernel_scalar_binding(body, written):
    """A scalar parameter the body never assigns (or takes the address of) is
    bound in kernels as a const copy, loop-invariant to the compiler; one it
    writes stays a reference to its device copy."""
    hipgen, _ = _loops()
    got = hipgen.scalar_bind(body, ('scalar', 'c', 'double'))
    assert got == ('double &c = *c_ptr;' if written else 'const double c = *c_ptr;'), got
