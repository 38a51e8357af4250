"""Device layout transforms and sub-array actualisation, compared with the
oracle bit for bit.

- ACROSS loops run with the TRANSFORM tuning forced on: line sweeps whose
  fastest free axis is not the array's last get the array's axes permuted
  (tile transpose / gather kernels), hyperplane wavefronts over three axes get
  the last two axes stored diagonal by diagonal; a following PARALLEL loop
  gets the natural layout back, and fetching a transformed array undoes the
  transform on a copy.
- get_actual / actual of sub-arrays (the block pack/unpack kernels), on plain
  arrays and on a diagonalised one; the host path (no GPU) runs the same
  program entry points on the CPU.
"""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

KIND_LINE, KIND_WAVEFRONT = 1, 2
T_TRANSFORM = 6
_c_lp = ctypes.POINTER(ctypes.c_long)
_c_dp = ctypes.POINTER(ctypes.c_double)


def _force(kind, block, tune):
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    if block is None:
        r.dvmh_amd_tuner_force_C(kind, None, None)
        return
    b = (ctypes.c_long * 3)(*block)
    t = (ctypes.c_long * 8)(*tune)
    r.dvmh_amd_tuner_force_C(kind, ctypes.cast(b, ctypes.c_void_p), ctypes.cast(t, ctypes.c_void_p))


def _layout_changes():
    r = D.runtime()
    r.dvmh_amd_layout_changes_C.restype = ctypes.c_long
    return r.dvmh_amd_layout_changes_C()


def _transform_tune():
    tune = [0] * 8
    tune[T_TRANSFORM] = 1
    return tune


def _bounds(lo, hi):
    return (ctypes.c_long * 3)(*lo), (ctypes.c_long * 3)(*hi)


def _block(a3, lo, hi):
    return a3[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]


# ---------------------------------------------------------------- transforms on the GPU
@pytest.mark.gpu
@pytest.mark.parametrize('block', [(32, 8, 1), (64, 4, 1), (256, 1, 1)])
@pytest.mark.parametrize('n,steps', [(6, 2), (33, 2), (64, 3)])
def test_adi3d_line_sweeps_permuted_layout(block, n, steps):
    """The k sweeps (dependent innermost axis) run over an array whose two
    last axes are swapped; the MAXLOC loop after them gets the natural layout
    back."""
    before = _layout_changes()
    _force(KIND_LINE, block, _transform_tune())
    try:
        res = H.run_adi3d(n, steps)
    finally:
        _force(KIND_LINE, None, None)
    changes = _layout_changes() - before
    ref = H.oracle_adi3d(n, steps)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])
    assert changes >= 2 * steps     # permuted for the k sweeps, natural again for MAXLOC


@pytest.mark.gpu
@pytest.mark.parametrize('block', [(64, 1, 1), (32, 2, 1), (128, 4, 1)])
@pytest.mark.parametrize('n,iters', [(5, 2), (23, 3), (70, 2)])
def test_sor3d_hyperplane_diagonal_layout(block, n, iters):
    """Hyperplane wavefronts over (j, k) planes stored diagonal by diagonal;
    the final fetch undoes the diagonal on a copy."""
    before = _layout_changes()
    _force(KIND_WAVEFRONT, block, _transform_tune())
    try:
        res = H.run_sor3d(n, iters)
    finally:
        _force(KIND_WAVEFRONT, None, None)
    changes = _layout_changes() - before
    ref = H.oracle_sor3d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['eps'], ref['eps'])
    assert changes >= 1


@pytest.mark.gpu
def test_adi3d_permuted_full_size():
    """256^3 (the ADI workload's size), one step, permuted k sweeps."""
    _force(KIND_LINE, (64, 4, 1), _transform_tune())
    try:
        res = H.run_adi3d(256, 1)
    finally:
        _force(KIND_LINE, None, None)
    ref = H.oracle_adi3d(256, 1)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['loc'], ref['loc'])


@pytest.mark.gpu
def test_sor3d_diagonal_full_size():
    """320^3 (the SOR workload's size), one sweep over diagonal planes."""
    _force(KIND_WAVEFRONT, (128, 2, 1), _transform_tune())
    try:
        res = H.run_sor3d(320, 1)
    finally:
        _force(KIND_WAVEFRONT, None, None)
    ref = H.oracle_sor3d(320, 1)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])


# ---------------------------------------------------------------- sub-arrays
BLOCKS = [((0, 0, 0), (0, 0, 0)), ((1, 2, 3), (4, 2, 7)), ((0, 5, 0), (10, 5, 10)), ((3, 0, 9), (3, 10, 10)),
          ((0, 0, 0), (10, 10, 10))]
N_SUB = 11


def _jac3d_get_actual_blocks(targets, iters=3):
    P = H._program('jac3d')
    P.set_targets(targets, 0)
    P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(N_SUB, D.SEED_BASE + 2)
    eps = np.zeros(iters)
    s = np.zeros(iters)
    P.fn('iterate', [ctypes.c_int, _c_dp, _c_dp], ctypes.c_double)(iters, H.dp(eps), H.dp(s))
    fetch_block = P.fn('fetch_block', [ctypes.c_int, _c_lp, _c_lp, _c_dp])
    got = []
    for lo, hi in BLOCKS:
        for which in (0, 1):
            ext = [h - l + 1 for l, h in zip(lo, hi)]
            out = np.zeros(int(np.prod(ext)))
            clo, chi = _bounds(lo, hi)
            fetch_block(which, clo, chi, H.dp(out))
            got.append(out.reshape(ext))
    P.fn('teardown', [])()
    ref = H.oracle_jac3d(N_SUB, iters)
    want = []
    for lo, hi in BLOCKS:
        for key in ('A', 'B'):
            want.append(_block(ref[key].reshape(N_SUB, N_SUB, N_SUB), lo, hi))
    return got, want


def test_jac3d_subarray_get_actual_host():
    got, want = _jac3d_get_actual_blocks(D.DEVICE_TYPE_HOST)
    for g, w in zip(got, want):
        assert H.bits_equal(g, w)


@pytest.mark.gpu
def test_jac3d_subarray_get_actual_gpu():
    """Blocks of a single element, a row, a plane, a column and the whole array,
    packed on the device and copied alone."""
    got, want = _jac3d_get_actual_blocks(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA)
    for g, w in zip(got, want):
        assert H.bits_equal(g, w)


def _sweep(b):
    """One 7-point Jacobi sweep of the interior, additions in program order."""
    a = b.copy()
    a[1:-1, 1:-1, 1:-1] = (b[:-2, 1:-1, 1:-1] + b[2:, 1:-1, 1:-1] + b[1:-1, :-2, 1:-1] + b[1:-1, 2:, 1:-1]
                           + b[1:-1, 1:-1, :-2] + b[1:-1, 1:-1, 2:] + b[1:-1, 1:-1, 1:-1]) / 7.0
    return a


def _jac3d_actual_block(targets, lo, hi, iters=3):
    """After three sweeps (B written last) a block of B is replaced on the host
    and made actual; the fourth sweep (B -> A) must read it."""
    P = H._program('jac3d')
    P.set_targets(targets, 0)
    P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(N_SUB, D.SEED_BASE + 2)
    eps = np.zeros(iters + 1)
    s = np.zeros(iters + 1)
    it = P.fn('iterate', [ctypes.c_int, _c_dp, _c_dp], ctypes.c_double)
    it(iters, H.dp(eps), H.dp(s))
    ext = [h - l + 1 for l, h in zip(lo, hi)]
    rng = np.random.default_rng(17)
    vals = rng.random(int(np.prod(ext)))
    clo, chi = _bounds(lo, hi)
    P.fn('store_block', [ctypes.c_int, _c_lp, _c_lp, _c_dp])(1, clo, chi, H.dp(vals))
    e1 = np.zeros(1)
    s1 = np.zeros(1)
    it(1, H.dp(e1), H.dp(s1))
    A = np.zeros(N_SUB ** 3)
    B = np.zeros(N_SUB ** 3)
    P.fn('fetch', [_c_dp, _c_dp])(H.dp(A), H.dp(B))
    P.fn('teardown', [])()
    ref = H.oracle_jac3d(N_SUB, iters)
    Bm = ref['B'].reshape(N_SUB, N_SUB, N_SUB).copy()
    Bm[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = vals.reshape(ext)
    Aw = _sweep(Bm)
    Aw[0, :, :] = Aw[-1, :, :] = 0.0
    Aw[:, 0, :] = Aw[:, -1, :] = 0.0
    Aw[:, :, 0] = Aw[:, :, -1] = 0.0
    return A.reshape(Aw.shape), B.reshape(Bm.shape), Aw, Bm


@pytest.mark.parametrize('lo,hi', BLOCKS[1:4])
def test_jac3d_subarray_actual_host(lo, hi):
    A, B, Aw, Bm = _jac3d_actual_block(D.DEVICE_TYPE_HOST, lo, hi)
    assert H.bits_equal(B, Bm) and H.bits_equal(A, Aw)


@pytest.mark.gpu
@pytest.mark.parametrize('lo,hi', BLOCKS)
def test_jac3d_subarray_actual_gpu(lo, hi):
    """The host writes a block of an array the device owns; only that block is
    sent, the rest of the device copy stays as the device wrote it."""
    A, B, Aw, Bm = _jac3d_actual_block(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, lo, hi)
    assert H.bits_equal(B, Bm) and H.bits_equal(A, Aw)


@pytest.mark.gpu
@pytest.mark.parametrize('lo,hi', [((1, 2, 3), (4, 2, 7)), ((0, 0, 0), (22, 22, 22)), ((5, 0, 0), (5, 22, 22))])
def test_sor3d_subarray_of_diagonal_array(lo, hi):
    """get_actual of a block of an array the last ACROSS loop left stored
    diagonal by diagonal, then the next sweeps and a full fetch."""
    n, iters = 23, 2
    _force(KIND_WAVEFRONT, (64, 1, 1), _transform_tune())
    try:
        P = H._program('sor3d')
        P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)
        P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(n, 1.5, D.SEED_BASE + 3)
        eps = np.zeros(iters)
        it = P.fn('iterate', [ctypes.c_int, _c_dp], ctypes.c_double)
        it(iters, H.dp(eps))
        ext = [h - l + 1 for l, h in zip(lo, hi)]
        out = np.zeros(int(np.prod(ext)))
        clo, chi = _bounds(lo, hi)
        P.fn('fetch_block', [_c_lp, _c_lp, _c_dp])(clo, chi, H.dp(out))
        eps2 = np.zeros(1)
        it(1, H.dp(eps2))
        A = np.zeros(n ** 3)
        P.fn('fetch', [_c_dp])(H.dp(A))
        P.fn('teardown', [])()
    finally:
        _force(KIND_WAVEFRONT, None, None)
    ref2 = H.oracle_sor3d(n, iters)
    ref3 = H.oracle_sor3d(n, iters + 1)
    assert H.bits_equal(out.reshape(ext), _block(ref2['A'].reshape(n, n, n), lo, hi))
    assert H.bits_equal(A, ref3['A'])
    assert np.array_equal(np.concatenate([eps, eps2]), ref3['eps'])
