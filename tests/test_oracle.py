"""Pins of the oracle (the CPU restatement of the reference's host handlers)
against what the mathematics fixes, and of the runtime's integer host logic
against the oracle -- so that a plausible mistake in either (a dropped stencil
term, a wrong neighbour index or sign, a wrong sweep order, a transposed
operand, a wrong tie rule) fails here.  Every check runs on the CPU."""
import ctypes

import numpy as np
import pytest
import scipy.linalg
import scipy.signal

import dvmh_amd as D
import helpers as H

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


# ---------------------------------------------------------------- input generator
def test_splitmix64_published_vectors():
    """SplitMix64 (Steele, Lea & Flood 2014) from state 0: the published first
    outputs; orc_mix64 is its output function applied to state + k * golden."""
    o = H.oracle()
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [o.orc_mix64((k * GOLDEN) & MASK64) for k in (1, 2, 3)]
    assert got == want


def test_u01_is_53_bit_uniform():
    o = H.oracle()
    vals = np.array([o.orc_u01(12345, n) for n in range(20000)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert np.all(vals * 2.0 ** 53 == np.floor(vals * 2.0 ** 53))   # multiples of 2^-53
    assert abs(vals.mean() - 0.5) < 0.01 and abs(vals.var() - 1 / 12) < 0.005


# ---------------------------------------------------------------- 2-D Jacobi (C1, C5)
def _jac2d(A, iters=1):
    n = int(round(np.sqrt(A.size)))
    A = A.copy()
    B = A.copy()
    eps = np.zeros(iters)
    H.oracle().orc_jac2d(H.vp(A), H.vp(B), n, iters, H.vp(eps), 0)
    return A, B, eps


def test_jac2d_leaves_a_linear_field_unchanged():
    """A linear function is discrete-harmonic: the 5-point average reproduces it
    exactly (integer values, exact in FP64), so eps = 0 and A is unchanged."""
    n = 33
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    A = (3.0 * i - 5.0 * j + 7.0).ravel()
    A2, B2, eps = _jac2d(A, 3)
    assert np.array_equal(A2, A) and np.all(eps == 0.0)


def test_jac2d_eigenmode_decay():
    """Zero boundary, A = sin(p pi i/(n-1)) sin(q pi j/(n-1)): one sweep scales
    the interior by (cos(p pi/(n-1)) + cos(q pi/(n-1))) / 2.  A dropped or
    shifted neighbour term changes the factor or breaks the mode shape."""
    n, p, q = 41, 3, 5
    x = np.arange(n) * np.pi / (n - 1)
    A = np.outer(np.sin(p * x), np.sin(q * x))
    A[0, :] = A[-1, :] = A[:, 0] = A[:, -1] = 0.0
    lam = (np.cos(p * np.pi / (n - 1)) + np.cos(q * np.pi / (n - 1))) / 2
    A1, B1, eps = _jac2d(A.ravel(), 1)
    B1 = B1.reshape(n, n)
    assert np.allclose(B1[1:-1, 1:-1], lam * A[1:-1, 1:-1], rtol=0, atol=1e-14)
    assert np.isclose(eps[0], np.max(np.abs((lam - 1) * A[1:-1, 1:-1])), rtol=1e-12)
    assert np.array_equal(A1.reshape(n, n)[1:-1, 1:-1], B1[1:-1, 1:-1])   # copy loop: A = B


# ---------------------------------------------------------------- 3-D Jacobi (C2)
def test_jac3d_eigenmode_decay_max_and_sum():
    """7-point average (six neighbours + centre) / 7 on sin modes: factor
    (2 cos a + 2 cos b + 2 cos c + 1) / 7; MAX = max |B - A|, SUM = sum of B."""
    n, modes = 19, (1, 2, 3)
    x = np.arange(n) * np.pi / (n - 1)
    s = [np.sin(m * x) for m in modes]
    A = np.einsum('i,j,k->ijk', *s)
    A[0] = A[-1] = 0
    A[:, 0] = A[:, -1] = 0
    A[:, :, 0] = A[:, :, -1] = 0
    lam = (sum(2 * np.cos(m * np.pi / (n - 1)) for m in modes) + 1) / 7
    a = A.ravel().copy()
    b = a.copy()
    eps = np.zeros(1)
    sm = np.zeros(1)
    H.oracle().orc_jac3d(H.vp(a), H.vp(b), n, 1, H.vp(eps), H.vp(sm), 0)
    B = b.reshape(n, n, n)
    inner = (slice(1, -1),) * 3
    assert np.allclose(B[inner], lam * A[inner], rtol=0, atol=1e-14)
    assert np.isclose(eps[0], np.max(np.abs(B[inner] - A[inner])), rtol=1e-12)
    assert np.isclose(sm[0], np.sum(B[inner]), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- SOR (C3)
def test_sor3d_sweep_equals_triangular_solve():
    """One lexicographic in-place SOR sweep of x = (w/6) * (sum of the six
    neighbours) + (1 - w) x is x' = (I - (w/6) L)^-1 ((w/6) U + (1 - w) I) x,
    L / U the adjacency of the lower (i-1, j-1, k-1) / upper neighbours.  A
    wrong dependency order (old instead of new values) breaks the equality."""
    n, w = 7, 1.5
    m = n - 2
    rng = np.random.default_rng(4)
    A = rng.uniform(size=(n, n, n))
    A[0] = A[-1] = 0
    A[:, 0] = A[:, -1] = 0
    A[:, :, 0] = A[:, :, -1] = 0
    idx = lambda i, j, k: ((i - 1) * m + (j - 1)) * m + (k - 1)
    N = m ** 3
    L = np.zeros((N, N))
    U = np.zeros((N, N))
    for i in range(1, n - 1):
        for j in range(1, n - 1):
            for k in range(1, n - 1):
                r = idx(i, j, k)
                for (di, dj, dk) in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    lo = (i - di, j - dj, k - dk)
                    hi = (i + di, j + dj, k + dk)
                    if min(lo) >= 1:
                        L[r, idx(*lo)] = 1
                    if max(hi) <= n - 2:
                        U[r, idx(*hi)] = 1
    x = A[1:-1, 1:-1, 1:-1].ravel()
    rhs = (w / 6) * (U @ x) + (1 - w) * x
    want = scipy.linalg.solve_triangular(np.eye(N) - (w / 6) * L, rhs, lower=True)
    a = A.ravel().copy()
    eps = np.zeros(1)
    H.oracle().orc_sor3d(H.vp(a), n, 1, w, H.vp(eps), 0)
    got = a.reshape(n, n, n)[1:-1, 1:-1, 1:-1].ravel()
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    assert np.isclose(eps[0], np.max(np.abs(want - x)), rtol=1e-12)


@pytest.mark.parametrize('threads', [1, 3, 8])
def test_sor3d_pipeline_independent_of_thread_count(threads):
    """The pipelined handshake gives the sequential order for any thread count."""
    n, w = 12, 1.3
    rng = np.random.default_rng(5)
    base = rng.uniform(size=n ** 3)
    ref = base.copy()
    e1 = np.zeros(2)
    H.oracle().orc_sor3d(H.vp(ref), n, 2, w, H.vp(e1), 1)
    got = base.copy()
    e2 = np.zeros(2)
    H.oracle().orc_sor3d(H.vp(got), n, 2, w, H.vp(e2), threads)
    assert H.bits_equal(got, ref) and np.array_equal(e1, e2)


# ---------------------------------------------------------------- ADI (C4)
def test_adi3d_step_equals_recursive_filters():
    """Forward sweep x[i] += c x[i-1] is the IIR filter y[i] = x[i] + c y[i-1];
    backward the same on the reversed axis.  One step = forward and backward
    along axis 0, 1, 2 in that order, then MAXLOC of |x| (lowest linear index
    among ties, as numpy.argmax)."""
    n, c = 9, 0.3
    rng = np.random.default_rng(6)
    A = rng.uniform(-1, 1, size=(n, n, n))
    want = A.copy()
    for ax in range(3):
        want = scipy.signal.lfilter([1.0], [1.0, -c], want, axis=ax)
        want = np.flip(scipy.signal.lfilter([1.0], [1.0, -c], np.flip(want, ax), axis=ax), ax)
    a = A.ravel().copy()
    mx = np.zeros(1)
    loc = np.zeros(1, dtype=np.int64)
    H.oracle().orc_adi3d(H.vp(a), n, 1, c, H.vp(mx), H.vp(loc), 0)
    assert np.allclose(a.reshape(n, n, n), want, rtol=1e-14, atol=1e-15)
    assert mx[0] == np.max(np.abs(a)) and loc[0] == int(np.argmax(np.abs(a)))


def test_adi3d_maxloc_tie_takes_lowest_index():
    """c = 0 leaves the array unchanged; equal maxima at several places give the
    lowest linear index (the sequential strict-'>' body)."""
    n = 10
    a = np.zeros(n ** 3)
    for p in (777, 123, 456):
        a[p] = -2.5
    mx = np.zeros(1)
    loc = np.zeros(1, dtype=np.int64)
    H.oracle().orc_adi3d(H.vp(a), n, 1, 0.0, H.vp(mx), H.vp(loc), 8)
    assert mx[0] == 2.5 and loc[0] == 123


# ---------------------------------------------------------------- diagonal layout index map
def _rt():
    r = D.runtime()
    r.dvmh_amd_convert_xy_C.restype = ctypes.c_long
    r.dvmh_amd_convert_xy_C.argtypes = [ctypes.c_long] * 5
    r.dvmh_change_filled_bounds2_C.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_long, ctypes.c_long, ctypes.c_void_p]
    return r


SIZES = [1, 2, 3, 5, 8, 13, 21, 34, 40]


@pytest.mark.parametrize('slash', [0, 1])
def test_convert_xy_is_a_diagonal_bijection(slash):
    """For every plane Rx x Ry: the map is a bijection onto [0, Rx*Ry); the
    points of each diagonal (x + y const, or x - y const with slash) occupy one
    contiguous run, runs in increasing order of that constant; the kernels' function (runtime
    export) and the oracle agree everywhere."""
    r = _rt()
    o = H.oracle()
    for Rx in SIZES:
        for Ry in SIZES:
            seen = np.zeros(Rx * Ry, dtype=bool)
            runs = {}
            for y in range(Ry):
                for x in range(Rx):
                    v = r.dvmh_amd_convert_xy_C(x, y, Rx, Ry, slash)
                    assert v == o.orc_convert_xy(x, y, Rx, Ry, slash)
                    assert 0 <= v < Rx * Ry and not seen[v], (Rx, Ry, x, y, v)
                    seen[v] = True
                    d = x + y if not slash else x - y
                    runs.setdefault(d, []).append(v)
            assert seen.all()
            prev_hi = -1
            for d in sorted(runs):
                lo, hi = min(runs[d]), max(runs[d])
                assert hi - lo + 1 == len(runs[d]) and lo == prev_hi + 1, (Rx, Ry, d)
                prev_hi = hi


def _diagonal_order_table(Rx, Ry, slash, flip=False):
    """Storage position of every (x, y) of an Rx x Ry plane, built by sorting
    rather than by a closed form.  The slash layout is the plain one with y
    mirrored (y -> Ry - 1 - y).  In the plain layout the cells are sorted by
    diagonal d = x + y; inside a diagonal they ascend with y when the plane is
    taller than wide (Rx < Ry), with x when it is wider (Rx > Ry), and for a
    square plane with y up to the main anti-diagonal (d < Rx) and with x after
    it.  This is the order rts/dvmh/include/dvmhlib_device.h:56-98 (and
    cdvmh/include/cdvmh_helpers.h:250-298) store diagonalised planes in.
    flip=True orders every diagonal by the other coordinate (a planted
    mistake: x and y exchanged inside a diagonal)."""
    cells = []
    for y in range(Ry):
        for x in range(Rx):
            ym = Ry - 1 - y if slash else y
            d = x + ym
            by_y = (Rx < Ry or (Rx == Ry and d < Rx)) != flip
            cells.append((d, ym if by_y else x, x, y))
    cells.sort()
    table = np.zeros((Ry, Rx), dtype=np.int64)
    for pos, (_, _, x, y) in enumerate(cells):
        table[y, x] = pos
    return table


@pytest.mark.parametrize('slash', [0, 1])
def test_convert_xy_order_inside_each_diagonal(slash):
    """The position inside a diagonal, not only the diagonal's run, is part of
    the diagonal storage format emitted handlers index: the runtime's function
    and the oracle equal the sorted table for every plane up to 40 x 40, and
    differ from the table with x and y exchanged inside the diagonals on every
    plane with a diagonal of two or more cells."""
    r = _rt()
    o = H.oracle()
    for Rx in SIZES:
        for Ry in SIZES:
            want = _diagonal_order_table(Rx, Ry, slash)
            for y in range(Ry):
                for x in range(Rx):
                    assert r.dvmh_amd_convert_xy_C(x, y, Rx, Ry, slash) == want[y, x], (Rx, Ry, x, y)
                    assert o.orc_convert_xy(x, y, Rx, Ry, slash) == want[y, x], (Rx, Ry, x, y)
            if min(Rx, Ry) > 1:
                got = np.array([[r.dvmh_amd_convert_xy_C(x, y, Rx, Ry, slash) for x in range(Rx)] for y in range(Ry)])
                assert not np.array_equal(got, _diagonal_order_table(Rx, Ry, slash, flip=True)), (Rx, Ry)


# ---------------------------------------------------------------- change_filled_bounds2
@pytest.mark.parametrize('rank', [1, 2, 3, 4])
def test_change_filled_bounds2_all_masks(rank):
    """rts/dvmh/dvmh2.cpp:4556-4583: dependent axes first (in axis order), then
    the others; bounds permuted accordingly when some but not all axes depend;
    0-based permutation out.  Runtime and oracle agree for every mask."""
    r = _rt()
    o = H.oracle()
    rng = np.random.default_rng(rank)
    for mask in range(1 << rank):
        lo = rng.integers(-5, 5, rank).astype(np.int64)
        hi = lo + rng.integers(0, 9, rank)
        st = rng.integers(1, 3, rank).astype(np.int64)
        outs = []
        for f in (r.dvmh_change_filled_bounds2_C, o.orc_change_filled_bounds2):
            a, b, c = lo.copy(), hi.copy(), st.copy()
            perm = np.zeros(rank, dtype=np.int64)
            f(H.vp(a), H.vp(b), H.vp(c), rank, mask, H.vp(perm))
            outs.append((a, b, c, perm))
        for x, y in zip(outs[0], outs[1]):
            assert np.array_equal(x, y), (rank, mask)
        a, b, c, perm = outs[0]
        dep = [i for i in range(rank) if (mask >> (rank - 1 - i)) & 1]
        assert list(perm) == dep + [i for i in range(rank) if i not in dep]
        if 0 < len(dep) < rank:
            assert np.array_equal(a, lo[perm]) and np.array_equal(b, hi[perm]) and np.array_equal(c, st[perm])
        else:
            assert np.array_equal(a, lo) and np.array_equal(b, hi) and np.array_equal(c, st)
