"""Shared test helpers: oracle bindings (CPU restatement, the checker) and
runners for the converted workload programs (the product path)."""
import ctypes
import os

import numpy as np

import dvmh_amd as D

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, 'oracle', '_build', 'liboracle.so')

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_lp = ctypes.POINTER(ctypes.c_long)
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError('oracle not built: run __graft_entry__.build()')
        o = ctypes.CDLL(ORACLE_SO)
        vp, L, I, U, Dd = ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        o.orc_mix64.restype = U
        o.orc_mix64.argtypes = [U]
        o.orc_u01.restype = Dd
        o.orc_u01.argtypes = [U, U]
        o.orc_init_grid2d.argtypes = [vp, L, U, Dd]
        o.orc_init_grid3d.argtypes = [vp, L, U, Dd, I]
        o.orc_batched_sides.argtypes = [vp, I, U]
        o.orc_batched_grid_seed.restype = U
        o.orc_batched_grid_seed.argtypes = [U, I]
        o.orc_jac2d.argtypes = [vp, vp, L, I, vp, I]
        o.orc_jac3d.argtypes = [vp, vp, L, I, vp, vp, I]
        o.orc_sor3d.argtypes = [vp, L, I, Dd, vp, I]
        o.orc_adi3d.argtypes = [vp, L, I, Dd, vp, vp, I]
        o.orc_adi_kat.restype = Dd
        o.orc_adi_kat.argtypes = [L, I, I, vp]
        o.orc_jac3d_f32_kat.restype = ctypes.c_float
        o.orc_jac3d_f32_kat.argtypes = [L, I, I, vp]
        o.orc_jac2d_f32_kat.restype = ctypes.c_float
        o.orc_jac2d_f32_kat.argtypes = [L, I, I, vp]
        o.orc_convert_xy.restype = L
        o.orc_convert_xy.argtypes = [L, L, L, L, I]
        o.orc_change_filled_bounds2.argtypes = [vp, vp, vp, L, L, vp]
        _oracle = o
    return _oracle


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dp(a):
    return None if a is None else a.ctypes.data_as(_c_dp)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


# ------------------------------------------------------------------ oracle runs
def oracle_jac2d(n, iters, seed=None, threads=0):
    seed = D.SEED_BASE + 1 if seed is None else seed
    o = oracle()
    A = np.zeros(n * n)
    o.orc_init_grid2d(vp(A), n, seed, 1.0)
    B = A.copy()
    eps = np.zeros(iters)
    o.orc_jac2d(vp(A), vp(B), n, iters, vp(eps), threads)
    return {'A': A, 'B': B, 'eps': eps}


# ------------------------------------------------------------------ program runs
def _program(name):
    P = D.Program(name)
    return P


def run_jac2d(n, iters, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0):
    seed = D.SEED_BASE + 1 if seed is None else seed
    P = _program('jac2d')
    P.set_targets(targets, flags)
    P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(n, seed)
    eps = np.zeros(iters)
    t = P.fn('iterate', [ctypes.c_int, _c_dp], ctypes.c_double)(iters, dp(eps))
    A = np.zeros(n * n)
    B = np.zeros(n * n)
    P.fn('fetch', [_c_dp, _c_dp])(dp(A), dp(B))
    P.fn('teardown', [])()
    return {'A': A, 'B': B, 'eps': eps, 'seconds': t}


def oracle_jac3d(n, iters, seed=None, threads=0):
    seed = D.SEED_BASE + 2 if seed is None else seed
    o = oracle()
    A = np.zeros(n ** 3)
    o.orc_init_grid3d(vp(A), n, seed, 0.0, 1)
    B = A.copy()
    eps = np.zeros(iters)
    s = np.zeros(iters)
    o.orc_jac3d(vp(A), vp(B), n, iters, vp(eps), vp(s), threads)
    return {'A': A, 'B': B, 'eps': eps, 'sum': s}


def oracle_sor3d(n, iters, w=1.5, seed=None, threads=0):
    seed = D.SEED_BASE + 3 if seed is None else seed
    o = oracle()
    A = np.zeros(n ** 3)
    o.orc_init_grid3d(vp(A), n, seed, 0.0, 1)
    eps = np.zeros(iters)
    o.orc_sor3d(vp(A), n, iters, w, vp(eps), threads)
    return {'A': A, 'eps': eps}


def oracle_adi3d(n, steps, c=0.3, seed=None, threads=0):
    seed = D.SEED_BASE + 4 if seed is None else seed
    o = oracle()
    A = np.zeros(n ** 3)
    o.orc_init_grid3d(vp(A), n, seed, 0.0, 0)
    mx = np.zeros(steps)
    loc = np.zeros(steps, dtype=np.int64)
    o.orc_adi3d(vp(A), n, steps, c, vp(mx), vp(loc), threads)
    return {'A': A, 'max': mx, 'loc': loc}


def run_jac3d(n, iters, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0, keep=False):
    seed = D.SEED_BASE + 2 if seed is None else seed
    P = _program('jac3d')
    P.set_targets(targets, flags)
    P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])(n, seed)
    eps = np.zeros(iters)
    s = np.zeros(iters)
    t = P.fn('iterate', [ctypes.c_int, _c_dp, _c_dp], ctypes.c_double)(iters, dp(eps), dp(s))
    A = np.zeros(n ** 3)
    B = np.zeros(n ** 3)
    P.fn('fetch', [_c_dp, _c_dp])(dp(A), dp(B))
    P.fn('teardown', [])()
    return {'A': A, 'B': B, 'eps': eps, 'sum': s, 'seconds': t}


def run_sor3d(n, iters, w=1.5, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0):
    seed = D.SEED_BASE + 3 if seed is None else seed
    P = _program('sor3d')
    P.set_targets(targets, flags)
    P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(n, w, seed)
    eps = np.zeros(iters)
    t = P.fn('iterate', [ctypes.c_int, _c_dp], ctypes.c_double)(iters, dp(eps))
    A = np.zeros(n ** 3)
    P.fn('fetch', [_c_dp])(dp(A))
    P.fn('teardown', [])()
    return {'A': A, 'eps': eps, 'seconds': t}


def run_sor3d_chain(n, iters, group, w=1.5, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0):
    """sor3d with `group` sweeps per batch (sor3d_iterate_chain): each group
    runs through the chain handler of the sweep loop."""
    seed = D.SEED_BASE + 3 if seed is None else seed
    P = _program('sor3d')
    P.set_targets(targets, flags)
    P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(n, w, seed)
    eps = np.zeros(iters)
    t = P.fn('iterate_chain', [ctypes.c_int, ctypes.c_int, _c_dp], ctypes.c_double)(iters, group, dp(eps))
    A = np.zeros(n ** 3)
    P.fn('fetch', [_c_dp])(dp(A))
    P.fn('teardown', [])()
    return {'A': A, 'eps': eps, 'seconds': t}


def run_adi3d(n, steps, c=0.3, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0):
    seed = D.SEED_BASE + 4 if seed is None else seed
    P = _program('adi3d')
    P.set_targets(targets, flags)
    P.fn('setup', [ctypes.c_long, ctypes.c_double, ctypes.c_ulonglong])(n, c, seed)
    mx = np.zeros(steps)
    loc = np.zeros(steps, dtype=np.int64)
    t = P.fn('iterate', [ctypes.c_int, _c_dp, _c_lp], ctypes.c_double)(steps, dp(mx), loc.ctypes.data_as(_c_lp))
    A = np.zeros(n ** 3)
    P.fn('fetch', [_c_dp])(dp(A))
    P.fn('teardown', [])()
    return {'A': A, 'max': mx, 'loc': loc, 'seconds': t}


def sum_close(got, ref, rtol=1e-12):
    """SUM reductions: within rtol relative of the compensated oracle sum."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    return bool(np.all(np.abs(got - ref) <= rtol * np.abs(ref)))


def run_adikat(n, itmax, maxeps=0.01, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA):
    P = _program('adikat')
    P.set_targets(targets, 0)
    eps = np.zeros(itmax)
    done = P.fn('run', [ctypes.c_long, ctypes.c_int, ctypes.c_double, _c_dp], ctypes.c_int)(n, itmax, maxeps, dp(eps))
    return eps[:done]


def run_jackat(rank, n, itmax, maxeps=0.5, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA):
    P = _program('jackat')
    P.set_targets(targets, 0)
    eps = np.zeros(itmax, dtype=np.float32)
    last = P.fn('run', [ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float)],
                ctypes.c_float)(rank, n, itmax, maxeps, eps.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return float(last), eps


# ------------------------------------------------------------------ batched uneven grids
def batched_sides(count, seed=None):
    seed = D.SEED_BASE + 5 if seed is None else seed
    s = np.zeros(count, dtype=np.int64)
    oracle().orc_batched_sides(vp(s), count, seed)
    return s


def oracle_batched(sides, iters, seed=None, grids=None, threads=0):
    """Each grid g: orc_jac2d on its own grid seeded by orc_batched_grid_seed."""
    seed = D.SEED_BASE + 5 if seed is None else seed
    o = oracle()
    out = []
    for g in (range(len(sides)) if grids is None else grids):
        n = int(sides[g])
        A = np.zeros(n * n)
        o.orc_init_grid2d(vp(A), n, o.orc_batched_grid_seed(seed, g), 1.0)
        B = A.copy()
        eps = np.zeros(iters)
        o.orc_jac2d(vp(A), vp(B), n, iters, vp(eps), threads)
        out.append({'g': g, 'A': A, 'B': B, 'eps': eps})
    return out


def run_batched(count, iters, batched, sides=None, seed=None, targets=D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, flags=0,
                fetch=True):
    """The converted batched program: returns eps history [iters, count], the
    grids (if fetch) and the iteration wall time."""
    seed = D.SEED_BASE + 5 if seed is None else seed
    P = _program('batched')
    P.set_targets(targets, flags)
    sides_arr = None if sides is None else np.ascontiguousarray(sides, dtype=np.int64)
    P.fn('setup', [ctypes.c_int, _c_lp, ctypes.c_ulonglong])(
        count, None if sides_arr is None else sides_arr.ctypes.data_as(_c_lp), seed)
    side = P.fn('side', [ctypes.c_int], ctypes.c_long)
    got_sides = np.array([side(g) for g in range(count)], dtype=np.int64)
    eps = np.zeros(iters * count)
    t = P.fn('iterate', [ctypes.c_int, ctypes.c_int, _c_dp], ctypes.c_double)(iters, int(batched), dp(eps))
    grids = []
    if fetch:
        f = P.fn('fetch', [ctypes.c_int, _c_dp, _c_dp])
        for g in range(count):
            n = int(got_sides[g])
            A = np.zeros(n * n)
            B = np.zeros(n * n)
            f(g, dp(A), dp(B))
            grids.append({'A': A, 'B': B})
    P.fn('teardown', [])()
    return {'sides': got_sides, 'eps': eps.reshape(iters, count), 'grids': grids, 'seconds': t}
