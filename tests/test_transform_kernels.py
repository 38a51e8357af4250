"""The layout-transform kernels called directly on device buffers
(csrc/kernels/transform.hip; what is moved where follows
rts/dvmh/cuda_transform.cu): dvmh_hip_permute against numpy.transpose for the
tiled transpose (the two fastest axes exchanged), the row gather (any other
permutation) and the byte gather of odd record sizes; dvmh_hip_diagonal
against the diagonal index map applied on the host element by element
(dvmh_amd_convert_xy_C, whose own ordering is pinned in test_oracle.py), and
back again.  Every byte must match."""
import ctypes
import itertools

import numpy as np
import pytest

import dvmh_amd as D


class Hip:
    def __init__(self):
        h = ctypes.CDLL('libamdhip64.so')
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h = h

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(nbytes, 16)) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def _lib():
    lib = D.runtime()
    lib.dvmh_hip_permute.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
    lib.dvmh_hip_diagonal.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                      ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.dvmh_amd_convert_xy_C.argtypes = [ctypes.c_long] * 5
    lib.dvmh_amd_convert_xy_C.restype = ctypes.c_long
    return lib


def _records(shape, rec, seed):
    """Seeded records of `rec` bytes: an array of shape + (rec,) uint8."""
    return np.random.default_rng(seed).integers(0, 256, size=tuple(shape) + (rec,), dtype=np.uint8)


def _permute(lib, hip, a, perm):
    rank = a.ndim - 1
    rec = a.shape[-1]
    sizes = (ctypes.c_long * rank)(*a.shape[:rank])
    p = (ctypes.c_int * rank)(*perm)
    src = hip.upload(np.ascontiguousarray(a))
    dst = hip.alloc(a.nbytes)
    try:
        lib.dvmh_hip_permute(src, dst, rec, rank, sizes, p, None)
        out = hip.download(dst, tuple(a.shape[q] for q in perm) + (rec,), np.uint8)
    finally:
        hip.free(src, dst)
    return out


SHAPES = [(5, 7, 9), (3, 64, 130), (2, 3, 4, 5), (1, 1, 300), (17, 1, 65)]


@pytest.mark.gpu
@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('rec', [1, 2, 3, 4, 8, 16])
def test_permute_matches_transpose(shape, rec):
    lib, hip = _lib(), Hip()
    rank = len(shape)
    a = _records(shape, rec, seed=rank * 100 + rec)
    perms = [p for p in itertools.permutations(range(rank)) if list(p) != list(range(rank))]
    if rank == 4:
        perms = perms[::5]
    for perm in perms:
        want = np.transpose(a, tuple(perm) + (rank,))
        got = _permute(lib, hip, a, perm)
        assert np.array_equal(got, want), (shape, rec, perm)


def _diag_map(lib, Rx, Ry, slash):
    m = np.empty((Ry, Rx), dtype=np.int64)
    for y in range(Ry):
        for x in range(Rx):
            m[y, x] = lib.dvmh_amd_convert_xy_C(x, y, Rx, Ry, slash)
    return m


@pytest.mark.gpu
@pytest.mark.parametrize('Rz,Ry,Rx', [(1, 1, 1), (2, 5, 5), (3, 7, 130), (2, 130, 7), (1, 64, 64), (2, 65, 129), (1, 200, 3)])
@pytest.mark.parametrize('slash', [0, 1])
@pytest.mark.parametrize('rec', [4, 8, 16])
def test_diagonal_matches_index_map(Rz, Ry, Rx, slash, rec):
    lib, hip = _lib(), Hip()
    a = _records((Rz, Ry, Rx), rec, seed=Rx * 7 + Ry * 3 + Rz + slash + rec)
    m = _diag_map(lib, Rx, Ry, slash)
    assert np.array_equal(np.sort(m.ravel()), np.arange(Rx * Ry))
    want = np.empty((Rz, Rx * Ry, rec), dtype=np.uint8)
    want[:, m.ravel(), :] = a.reshape(Rz, Rx * Ry, rec)
    src = hip.upload(a)
    mid = hip.alloc(a.nbytes)
    back = hip.alloc(a.nbytes)
    try:
        lib.dvmh_hip_diagonal(src, mid, rec, Rx, Ry, Rz, 0, slash, None)
        got = hip.download(mid, want.shape, np.uint8)
        lib.dvmh_hip_diagonal(mid, back, rec, Rx, Ry, Rz, 1, slash, None)
        again = hip.download(back, a.shape, np.uint8)
    finally:
        hip.free(src, mid, back)
    assert np.array_equal(got, want)
    assert np.array_equal(again, a)


@pytest.mark.gpu
def test_layout_transform_timing():
    lib = D.runtime()
    lib.dvmh_amd_layout_transform_seconds_C.argtypes = [ctypes.c_long] * 5 + [ctypes.POINTER(ctypes.c_double)]
    lib.dvmh_amd_layout_transform_seconds_C.restype = ctypes.c_long
    out = (ctypes.c_double * 2)()
    for which in range(3):
        assert lib.dvmh_amd_layout_transform_seconds_C(which, 16, 64, 96, 2, out) == 0
        assert 0 < out[0] < 1 and 0 < out[1] < 1
    assert lib.dvmh_amd_layout_transform_seconds_C(3, 16, 64, 96, 2, out) == -1
