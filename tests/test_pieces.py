"""Actual-state tracking by pieces (rts/dvmh/dvmh_pieces.cpp and the actual
state of rts/dvmh/dvmh_data.cpp:339-600): where each copy of an array is
current is a set of disjoint boxes.  The box algebra is checked on the CPU
against boolean masks (random sequences of unions and differences, ranks 1-4,
the parts of a query box the set misses); on the GPU, programs/pieces.cdv
moves sub-arrays between host and device and the transfer counters show that
only the stale pieces move, with the results equal to a sequential
evaluation."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

_lp = ctypes.POINTER(ctypes.c_long)


def _eval():
    f = D.runtime().dvmh_amd_pieces_eval_C
    f.argtypes = [ctypes.c_long, _lp, ctypes.c_long, _lp, _lp, ctypes.c_long]
    f.restype = ctypes.c_long
    return f


def _mask_of(boxes, shape):
    m = np.zeros(shape, dtype=bool)
    for lo, hi in boxes:
        m[tuple(slice(l, h + 1) for l, h in zip(lo, hi))] = True
    return m


@pytest.mark.parametrize('rank', [1, 2, 3, 4])
def test_piece_algebra_matches_masks(rank):
    """Random unions and differences of boxes (some empty, some outside the
    grid's middle), then the missing parts of a random query: the boxes the
    runtime returns are disjoint and cover exactly query minus the set."""
    f = _eval()
    rng = np.random.default_rng(rank)
    n = [7, 6, 5, 4][:rank]
    for trial in range(60):
        ops = []
        mask = np.zeros(n, dtype=bool)
        for _ in range(rng.integers(1, 9)):
            lo = [int(rng.integers(0, s)) for s in n]
            hi = [int(min(s - 1, l + rng.integers(-1, s))) for l, s in zip(lo, n)]   # hi < lo: an empty box
            code = 1 if rng.random() < 0.65 else 2
            ops.append([code] + lo + hi)
            sl = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
            mask[sl] = code == 1
        qlo = [int(rng.integers(0, s)) for s in n]
        qhi = [int(rng.integers(l, s)) for l, s in zip(qlo, n)]
        opa = np.array(ops, dtype=np.int64).ravel()
        q = np.array(qlo + qhi, dtype=np.int64)
        out = np.zeros(4096 * 2 * rank, dtype=np.int64)
        cnt = f(rank, opa.ctypes.data_as(_lp), len(ops), q.ctypes.data_as(_lp), out.ctypes.data_as(_lp), 4096)
        boxes = [(out[i * 2 * rank:i * 2 * rank + rank], out[i * 2 * rank + rank:(i + 1) * 2 * rank]) for i in range(cnt)]
        got = np.zeros(n, dtype=np.int64)
        for lo, hi in boxes:
            assert all(l <= h for l, h in zip(lo, hi)), (trial, boxes)
            got[tuple(slice(l, h + 1) for l, h in zip(lo, hi))] += 1
        assert got.max(initial=0) <= 1, ('overlapping boxes', trial)
        want = _mask_of([(qlo, qhi)], n) & ~mask
        assert np.array_equal(got.astype(bool), want), trial


def _lib():
    return D.load('libcdv_pieces.so')


def pieces_reference(n):
    """Sequential execution of programs/pieces.cdv."""
    A = np.add.outer(np.arange(n) * 1000.0, np.arange(n) * 1.0)
    A[1, :] = A[1, :] * 2.0 + A[2, :]       # host rewrites row 1 after get_actual of rows 1-2
    A += 1.0                                 # region on the device
    A[:, 3] = -A[:, 3]                       # host rewrites column 3 after get_actual of the whole array
    A = A * 0.5                              # region on the device
    return A.ravel()


def _run(n):
    f = _lib().pieces
    f.argtypes = [ctypes.c_long, ctypes.POINTER(ctypes.c_double), _lp]
    out = np.zeros(n * n)
    moved = np.zeros(8, dtype=np.int64)
    f(n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), moved.ctypes.data_as(_lp))
    return out, moved


def test_pieces_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    out, _ = _run(9)
    assert H.bits_equal(out, pieces_reference(9))


@pytest.mark.gpu
@pytest.mark.parametrize('n', [6, 64, 300])
def test_pieces_move_only_stale_blocks(n):
    out, moved = _run(n)
    assert H.bits_equal(out, pieces_reference(n))
    line, whole = n * 8, n * n * 8
    # moved[k]: bytes moved in either direction across step k of programs/pieces.cdv
    assert moved[0] == 2 * line       # get_actual of rows 1-2: those rows come down
    assert moved[1] == line           # region entry: only row 1, rewritten on the host, goes up
    assert moved[2] == whole          # get_actual of the array the region wrote
    assert moved[3] == 0              # again: the host copy is current, nothing moves
    assert moved[4] == line           # region entry: only column 3, rewritten on the host, goes up
    assert moved[5] == whole
