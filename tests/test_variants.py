"""Kernel variants the tuner can choose, forced one at a time and compared with
the oracle bit for bit: the LDS march variant of PARALLEL stencil loops
(rolling three-plane window, several tile shapes, points per thread and march
lengths, including tiles that overhang the loop on every axis), the tiled
wavefront of ACROSS loops on three axes, and the tiled line sweep of ACROSS
loops along the innermost axis (forward and backward, chunks that do not
divide the line, a single-iteration chunk)."""
import ctypes

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

KIND_STENCIL = 3
T_DEPTH, T_XCD, T_VARIANT, T_CHUNK, T_TILE_A = 0, 1, 2, 3, 4


def _force(kind, block, tune):
    r = D.runtime()
    r.dvmh_amd_tuner_force_C.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    if block is None:
        r.dvmh_amd_tuner_force_C(kind, None, None)
        return
    b = (ctypes.c_long * 3)(*block)
    t = (ctypes.c_long * 8)(*tune)
    r.dvmh_amd_tuner_force_C(kind, ctypes.cast(b, ctypes.c_void_p), ctypes.cast(t, ctypes.c_void_p))


def _march(bx, by, cols, march, xcd=0):
    tune = [0] * 8
    tune[T_DEPTH] = 1
    tune[T_VARIANT] = 1
    tune[T_TILE_A] = cols
    tune[T_CHUNK] = march
    tune[T_XCD] = xcd
    return (bx, by, 1), tune


MARCH3 = [(64, 1, 1, 8, 0), (128, 4, 3, 16, 1), (256, 2, 2, 32, 0), (64, 8, 4, 8, 1), (128, 2, 1, 64, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', MARCH3)
@pytest.mark.parametrize('n,iters', [(10, 3), (66, 4), (131, 3)])
def test_jac3d_march_variant_bit_exact(cfg, n, iters):
    block, tune = _march(*cfg)
    _force(KIND_STENCIL, block, tune)
    try:
        res = H.run_jac3d(n, iters)
    finally:
        _force(KIND_STENCIL, None, None)
    ref = H.oracle_jac3d(n, iters)
    assert H.bits_equal(res['A'], ref['A']) and H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])
    assert H.sum_close(res['sum'], ref['sum'])


MARCH2 = [(64, 1, 1, 8, 0), (256, 1, 4, 32, 1), (1024, 1, 2, 16, 0), (128, 1, 3, 64, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', MARCH2)
@pytest.mark.parametrize('n,iters', [(5, 3), (70, 5), (1000, 3)])
def test_jac2d_march_variant_bit_exact(cfg, n, iters):
    block, tune = _march(*cfg)
    _force(KIND_STENCIL, block, tune)
    try:
        res = H.run_jac2d(n, iters)
    finally:
        _force(KIND_STENCIL, None, None)
    ref = H.oracle_jac2d(n, iters)
    assert H.bits_equal(res['A'], ref['A']) and H.bits_equal(res['B'], ref['B'])
    assert np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
def test_march_variant_full_size_jac3d():
    """384^3 x 4 iterations with the variant's likely-best shape."""
    block, tune = _march(128, 2, 3, 16, 1)
    _force(KIND_STENCIL, block, tune)
    try:
        res = H.run_jac3d(384, 4)
    finally:
        _force(KIND_STENCIL, None, None)
    ref = H.oracle_jac3d(384, 4)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])
    assert H.sum_close(res['sum'], ref['sum'])


# ---------------------------------------------------------------- tiled wavefront (ACROSS on three axes)
KIND_WAVEFRONT = 2


def _tile(nt, ta, tc):
    tune = [0] * 8
    tune[T_DEPTH] = 1
    tune[T_VARIANT] = 1
    tune[T_TILE_A] = ta
    tune[5] = tc          # TILE_B
    return (nt, 1, 1), tune


TILES = [(64, 4, 8), (256, 8, 16), (512, 16, 16), (128, 8, 64), (256, 12, 32)]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', TILES)
@pytest.mark.parametrize('n,iters', [(5, 2), (23, 3), (70, 2)])
def test_sor3d_tiled_wavefront_bit_exact(cfg, n, iters):
    """Tiles that overhang the loop on every axis, and tiles larger than the
    whole loop (n = 5: 3 x 3 x 3 iterations)."""
    block, tune = _tile(*cfg)
    _force(KIND_WAVEFRONT, block, tune)
    try:
        res = H.run_sor3d(n, iters)
    finally:
        _force(KIND_WAVEFRONT, None, None)
    ref = H.oracle_sor3d(n, iters)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['eps'], ref['eps'])


@pytest.mark.gpu
def test_sor3d_tiled_full_size():
    """320^3 (the SOR workload's size), 2 sweeps, against the oracle."""
    block, tune = _tile(256, 16, 16)
    _force(KIND_WAVEFRONT, block, tune)
    try:
        res = H.run_sor3d(320, 2)
    finally:
        _force(KIND_WAVEFRONT, None, None)
    ref = H.oracle_sor3d(320, 2)
    assert H.bits_equal(res['A'], ref['A']) and np.array_equal(res['eps'], ref['eps'])


# ---------------------------------------------------------------- tiled line sweep (ACROSS along the innermost axis)
KIND_LINE = 1


def _ltile(tl, ck, reg=0):
    """reg 0: LDS chunks; 1: the walk on a register copy of the LDS chunk;
    2: the streamed sweep (variant 3: a wave's lines in aligned 128-byte
    chunks through LDS; ck is not used)."""
    tune = [0] * 8
    tune[T_VARIANT] = 3 if reg == 2 else 1
    tune[T_TILE_A] = ck
    tune[5] = 1 if reg == 1 else 0   # TILE_B
    return (tl, 1, 1), tune


LTILES = [(64, 8, 0), (128, 16, 0), (256, 8, 0), (64, 32, 0), (128, 8, 0), (64, 8, 1), (128, 16, 1), (256, 32, 1),
          (64, 16, 2), (128, 32, 2), (256, 16, 2)]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', LTILES)
@pytest.mark.parametrize('n,steps', [(6, 2), (33, 2), (64, 3), (100, 2)])
def test_adi3d_tiled_line_sweep_bit_exact(cfg, n, steps):
    """The two sweeps along k use the tiled variant (the other four fall back
    to the direct line kernel with the same block); n = 6 gives one partial
    chunk, n = 33 and 64 several with a partial last one."""
    block, tune = _ltile(*cfg)
    _force(KIND_LINE, block, tune)
    try:
        res = H.run_adi3d(n, steps)
    finally:
        _force(KIND_LINE, None, None)
    ref = H.oracle_adi3d(n, steps)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])


@pytest.mark.gpu
@pytest.mark.parametrize('tl', [64, 256])
@pytest.mark.parametrize('n,its', [(45, 4), (130, 3), (384, 2)])
def test_adikat_streamed_line_sweep_bit_exact(tl, n, its):
    """The reference ADI's k-sweep reads both neighbours (the one ahead not yet
    updated) and folds a MAX: the streamed sweep with rows of every alignment
    (n = 45, 130) and whole 128-byte chunks (384), eps of every iteration equal
    to the oracle's."""
    block, tune = _ltile(tl, 16, 2)
    _force(KIND_LINE, block, tune)
    try:
        eps = H.run_adikat(n, its, maxeps=0.0)
    finally:
        _force(KIND_LINE, None, None)
    ref = np.zeros(its)
    H.oracle().orc_adi_kat(n, its, 0, H.vp(ref))
    assert np.array_equal(eps, ref), (eps, ref)


def _lreg(bx, by, ck, xcd=0):
    tune = [0] * 8
    tune[T_VARIANT] = 2
    tune[T_TILE_A] = ck
    tune[T_XCD] = xcd
    return (bx, by, 1), tune


LREGS = [(64, 1, 8, 0), (128, 1, 16, 1), (256, 1, 8, 0), (32, 4, 16, 0), (64, 2, 8, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', LREGS)
@pytest.mark.parametrize('n,steps', [(3, 2), (6, 2), (19, 2), (64, 3)])
def test_adi3d_register_line_sweep_bit_exact(cfg, n, steps):
    """The four sweeps along i and j (forward and backward) use the
    register-staged variant (the two along k fall back to the direct line
    kernel with the same block); n = 3 and 6 give one partial chunk, 19 and 64
    several with a partial last one for CK = 8 and 16."""
    block, tune = _lreg(*cfg)
    _force(KIND_LINE, block, tune)
    try:
        res = H.run_adi3d(n, steps)
    finally:
        _force(KIND_LINE, None, None)
    ref = H.oracle_adi3d(n, steps)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])


@pytest.mark.gpu
def test_adi3d_register_line_sweep_full_size():
    """256^3 (the ADI workload's size), two steps, register-staged i and j sweeps."""
    block, tune = _lreg(64, 1, 16)
    _force(KIND_LINE, block, tune)
    try:
        res = H.run_adi3d(256, 2)
    finally:
        _force(KIND_LINE, None, None)
    ref = H.oracle_adi3d(256, 2)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])


@pytest.mark.gpu
@pytest.mark.parametrize('cfg', [(128, 16, 1), (64, 16, 2)])
def test_adi3d_tiled_line_sweep_full_size(cfg):
    """256^3 (the ADI workload's size), two steps: the walk on a register copy
    of LDS chunks, and the streamed sweep."""
    block, tune = _ltile(*cfg)
    _force(KIND_LINE, block, tune)
    try:
        res = H.run_adi3d(256, 2)
    finally:
        _force(KIND_LINE, None, None)
    ref = H.oracle_adi3d(256, 2)
    assert H.bits_equal(res['A'], ref['A'])
    assert np.array_equal(res['max'], ref['max']) and np.array_equal(res['loc'], ref['loc'])
