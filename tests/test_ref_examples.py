"""The reference's own C-DVMH example programs (examples/jac2d.cdv,
examples/jacross.cdv, examples/jac3d.cdv, examples/sor.cdv, examples/redb_c.cdv
-- FP32, a function-local distributed array released before `return`,
examples/adi.cdv -- a distributed array passed to a function that inherits it,
the FP32 benchmarks examples/jac3d_perf.cdv and examples/jac2d_perf.cdv, and
examples/testcdvm.cdv -- an ordinary C array in a parallel loop outside regions),
read from the reference tree,
converted by the front end (emit/cdvmh.py: file-scope distributed arrays and
variables, the programs' macros, get_actual and fwrite of whole arrays),
their handlers emitted (emit/hipgen.py), compiled, linked against
libdvmh_amd and run.  What each prints (the per-iteration MAX reduction) and
the array it writes to jacobi.dat must equal a numpy restatement of the
program, bit for bit; the benchmark programs' own verification lines must
read SUCCESSFUL.

These run where the reference tree is present (on the host path when no GPU
is visible); the programs and their conversions are never stored in the
repository.  Without the reference tree the module is skipped.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

REF = '/root/reference/dvm_sys/examples'
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'dvm-snapshot-r6603-fork_amd')
REPO = os.path.dirname(PKG)
sys.path.insert(0, os.path.join(REPO, 'oracle'))
import build_ref_examples as BRE  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason='the reference tree is not present')


def build_and_run(stem, tmp_path, sizes=None):
    """sizes: {macro: value} -- the program's `#define macro N` lines are
    rewritten to N in a copy under tmp_path before conversion."""
    out = str(tmp_path)
    exe = os.path.join(out, stem)
    BRE.build(stem, exe, out, sizes)
    return run_example(exe, out)


def run_example(exe, cwd):
    """Runs a built example in `cwd`: its stdout and the jacobi.dat it writes (if any)."""
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    dat = os.path.join(cwd, 'jacobi.dat')
    return r.stdout, (np.fromfile(dat) if os.path.exists(dat) else None)


def eps_lines(hist):
    return ''.join('it=%4i   eps=%e\n' % (it, e) for it, e in hist)


def ref_jac2d(L=8, itmax=20, maxeps=0.5):
    """examples/jac2d.cdv: A = 0, B = 3 + i + j inside a zero boundary; per
    iteration eps = max |B - A| with A = B, then the 4-point mean of A into B."""
    A = np.zeros((L, L))
    B = np.zeros((L, L))
    for i in range(1, L - 1):
        for j in range(1, L - 1):
            B[i, j] = 3 + i + j
    hist = []
    for it in range(1, itmax + 1):
        eps = 0.0
        for i in range(1, L - 1):
            for j in range(1, L - 1):
                eps = max(abs(B[i, j] - A[i, j]), eps)
                A[i, j] = B[i, j]
        for i in range(1, L - 1):
            for j in range(1, L - 1):
                B[i, j] = (A[i - 1, j] + A[i, j - 1] + A[i, j + 1] + A[i + 1, j]) / 4.0
        hist.append((it, eps))
        if eps < maxeps:
            break
    return hist, B.ravel()


def ref_jacross(L=8, itmax=20, maxeps=0.5):
    """examples/jacross.cdv: in-place lexicographic sweeps (ACROSS on both
    axes) of the 4-point mean, eps = max |new - old|."""
    A = np.zeros((L, L))
    for i in range(1, L - 1):
        for j in range(1, L - 1):
            A[i, j] = 3.0 + i + j
    hist = []
    for it in range(1, itmax + 1):
        eps = 0.0
        for i in range(1, L - 1):
            for j in range(1, L - 1):
                b = (A[i - 1, j] + A[i + 1, j] + A[i, j - 1] + A[i, j + 1]) / 4.0
                eps = max(abs(b - A[i, j]), eps)
                A[i, j] = b
        hist.append((it, eps))
        if eps < maxeps:
            break
    return hist, A.ravel()


def ref_jac3d(L=8, itmax=20, maxeps=0.5):
    """examples/jac3d.cdv: the 3-D version (B = 4 + i + j + k inside, 6-point mean)."""
    A = np.zeros((L, L, L))
    B = np.zeros((L, L, L))
    for i in range(1, L - 1):
        for j in range(1, L - 1):
            for k in range(1, L - 1):
                B[i, j, k] = 4 + i + j + k
    hist = []
    for it in range(1, itmax + 1):
        eps = 0.0
        for i in range(1, L - 1):
            for j in range(1, L - 1):
                for k in range(1, L - 1):
                    eps = max(abs(B[i, j, k] - A[i, j, k]), eps)
                    A[i, j, k] = B[i, j, k]
        for i in range(1, L - 1):
            for j in range(1, L - 1):
                for k in range(1, L - 1):
                    B[i, j, k] = (A[i - 1, j, k] + A[i, j - 1, k] + A[i, j, k - 1] + A[i, j, k + 1] + A[i, j + 1, k] +
                                  A[i + 1, j, k]) / 6.0
        hist.append((it, eps))
        if eps < maxeps:
            break
    return hist, B.ravel()


def _relax_point(A, i, j, w):
    """(w / 4) * (4-point sum) + (1 - w) * A[i][j] in FP32, the programs' order."""
    f = np.float32
    return (w / f(4)) * (((A[i - 1, j] + A[i + 1, j]) + A[i, j - 1]) + A[i, j + 1]) + (f(1) - w) * A[i, j]


def _fp32_start(n):
    A = np.full((n, n), np.float32(-1.0), dtype=np.float32)
    for i in range(n):
        A[i, i] = np.float32(n + 2)
    return A


def ref_sor(n=10, itmax=20, maxeps=np.float32(0.5e-5), w=np.float32(0.5)):
    """examples/sor.cdv: FP32 in-place lexicographic over-relaxation (ACROSS on
    both axes), eps = max fabs(old - new) folded into an FP32 variable."""
    A = _fp32_start(n)
    hist = []
    for it in range(1, itmax + 1):
        eps = np.float32(0.0)
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                s = A[i, j]
                A[i, j] = _relax_point(A, i, j, w)
                d = abs(float(np.float32(s - A[i, j])))
                eps = np.float32(d) if d > float(eps) else eps
        hist.append((it, float(eps)))
        if eps < maxeps:
            break
    return hist, None


def ref_redb(n=8, itmax=10, maxeps=np.float32(0.5e-5), w=np.float32(0.5)):
    """examples/redb_c.cdv: FP32 red points (i + j odd, MAX of the change),
    then black points, each a PARALLEL loop."""
    A = _fp32_start(n)
    hist = []
    for it in range(1, itmax + 1):
        eps = np.float32(0.0)
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                if (i + j) % 2 == 1:
                    s = A[i, j]
                    A[i, j] = _relax_point(A, i, j, w)
                    d = abs(float(np.float32(s - A[i, j])))
                    eps = np.float32(d) if d > float(eps) else eps
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                if (i + j) % 2 == 0:
                    A[i, j] = _relax_point(A, i, j, w)
        hist.append((it, float(eps)))
        if eps < maxeps:
            break
    return hist, None


def ref_adi(n, itmax=100, maxeps=0.01):
    """examples/adi.cdv (the array reaches its initialisation through an
    inherited parameter): boundary 10 i/(n-1) + 10 j/(n-1) + 10 k/(n-1), zero
    inside; per iteration an in-place sweep along i, then along j, of the mean
    of the two neighbours on that axis, then along k with eps = max |a - mean|."""
    a = np.zeros((n, n, n))
    r = np.arange(n, dtype=np.float64)
    edge = 10.0 * r / (n - 1)
    full = edge[:, None, None] + edge[None, :, None] + edge[None, None, :]
    inner = np.zeros((n, n, n), dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    a[~inner] = full[~inner]
    lines = []
    for it in range(1, itmax + 1):
        for i in range(1, n - 1):
            a[i, 1:-1, 1:-1] = (a[i - 1, 1:-1, 1:-1] + a[i + 1, 1:-1, 1:-1]) / 2
        for j in range(1, n - 1):
            a[1:-1, j, 1:-1] = (a[1:-1, j - 1, 1:-1] + a[1:-1, j + 1, 1:-1]) / 2
        eps = 0.0
        for k in range(1, n - 1):
            t = (a[1:-1, 1:-1, k - 1] + a[1:-1, 1:-1, k + 1]) / 2
            eps = max(eps, float(np.abs(a[1:-1, 1:-1, k] - t).max()))
            a[1:-1, 1:-1, k] = t
        lines.append(' IT = %4i   EPS = %14.7E\n' % (it, eps))
        if eps < maxeps:
            break
    return ''.join(lines), eps


def test_reference_adi(tmp_path):
    """examples/adi.cdv at 20^3: every iteration's printed eps equals the
    restatement's; the timing line aside, the closing report follows."""
    out, _ = build_and_run('adi', tmp_path, sizes={'nx': 20, 'ny': 20, 'nz': 20})
    lines, eps = ref_adi(20)
    assert out.startswith(lines)
    tail = out[len(lines):].splitlines()
    assert tail[0] == ' ADI Benchmark Completed.' and tail[1] == ' Size            =   20 x   20 x   20'
    assert tail[-2] == ' Verification    =       %12s' % ('SUCCESSFUL' if abs(eps - 0.07249074) < 1e-6 else 'UNSUCCESSFUL')


def test_reference_adi_full_size(tmp_path):
    """examples/adi.cdv as written (384^3, 100 iterations, about 10 s on the
    host path): the program's own check |eps - 0.07249074| < 1e-6 passes."""
    out, _ = build_and_run('adi', tmp_path)
    assert ' IT =  100   EPS =  7.2490737E-02\n' in out
    assert ' Verification    =         SUCCESSFUL\n' in out


def test_reference_testcdvm(tmp_path):
    """examples/testcdvm.cdv: a region fills A and the 8-neighbour sum into B;
    a parallel loop outside regions (on the host) compares B with the same sum
    over an ordinary C array and MIN-reduces the first mismatch; the program
    reports complete when there is none."""
    out, _ = build_and_run('testcdvm', tmp_path)
    assert out == 'T_CDVM  -  complete\n'


@pytest.mark.parametrize('stem,last', [('jac3d_perf', ' IT =  100   EPS =  5.0580444E+00\n'),
                                       ('jac2d_perf', ' IT =  100   EPS =  5.8375977E+01\n')])
def test_reference_benchmark_checks(stem, last, tmp_path):
    """examples/jac3d_perf.cdv (FP32 384^3) and examples/jac2d_perf.cdv (FP32
    8000^2), 100 iterations each as written: the programs' own checks
    (|eps - 5.058044| < 1e-4, |eps - 58.37598| < 1e-3) pass."""
    out, _ = build_and_run(stem, tmp_path)
    assert last in out
    assert ' Verification    =         SUCCESSFUL\n' in out


@pytest.mark.parametrize('stem,ref', [('jac2d', ref_jac2d), ('jacross', ref_jacross), ('jac3d', ref_jac3d),
                                      ('sor', ref_sor), ('redb_c', ref_redb)])
def test_reference_example_program(stem, ref, tmp_path):
    out, data = build_and_run(stem, tmp_path)
    hist, final = ref()
    assert out == eps_lines(hist)
    if final is not None:
        assert data.shape == final.shape and np.array_equal(data.view(np.uint64), final.view(np.uint64))


@pytest.mark.parametrize('stem', ['gauss_c', 'gauss_wb'])
def test_gauss_examples(stem, tmp_path):
    """examples/gauss_c.cdv and examples/gauss_wb.cdv: arrays declared without
    a distribution, redistribute with WGTBLOCK weights (gauss_wb), realign,
    remote_access in a loop clause and on a host block, a region over
    sub-arrays and sequential parts of regions.  The system is the identity
    with a right-hand side of ones, so every X[j] the program prints is 1."""
    out, _ = build_and_run(stem, tmp_path)
    assert out == ''.join('j=%4i   X[j]=%e\n' % (j, 1.0) for j in range(8, -1, -1))


def test_mgrid_old_directive_form(tmp_path):
    """examples/mgrid.cdv is written in the old C-DVM form (DVM(...) macros the
    program defines as nothing, DO(...) loop-header macros), which cdvmh
    leaves to the preprocessor: the program converts and runs as plain C.
    Its output is the golden one of examples/democ_res.uni:323-324."""
    out, _ = build_and_run('mgrid', tmp_path)
    assert out == 'N1=8 M1=12 K1=16 \nN1=4 M1=6 K1=8 \n'


def tsk_lp_reference():
    """numpy restatement of examples/tsk_lp.cdv (FP32 grid blocks of two tasks,
    boundary exchange between the blocks, then a 4-point relaxation and
    MAX(|B - A|) per block; K = 8, Nt = {4, 4}, 10 iterations)."""
    K, Nt, ntask = 8, [4, 4], 2
    f32 = np.float32
    At, Bt = [], []
    for t in range(ntask):
        N, N0 = Nt[t], sum(Nt[i] - 1 for i in range(t))
        A = np.zeros((N + 1, K), dtype=f32)
        for i in range(N + 1):
            for j in range(K):
                edge = j == 0 or j == K - 1 or (t == 0 and i == 0) or (t == ntask - 1 and i == N)
                A[i, j] = f32(0) if edge else f32(1 + i + N0 + j)
        At.append(A)
        Bt.append(A.copy())
    lines = ['Task %d mapped onto processors 0..0' % t for t in range(ntask)] + ['-' * 37]
    for it in range(1, 11):
        eps0 = 0.0
        for t in range(ntask):
            A = At[t]
            if t != 0:
                A[0, :] = Bt[t - 1][Nt[t - 1] - 1, :]
            if t != ntask - 1:
                A[Nt[t], :] = Bt[t + 1][1, :]
        for t in range(ntask):
            A, B, N = At[t], Bt[t], Nt[t]
            eps = 0.0
            for i in range(1, N):
                for j in range(1, K - 1):
                    B[i, j] = (((A[i - 1, j] + A[i, j - 1]) + A[i + 1, j]) + A[i, j + 1]) / f32(4)
            for i in range(1, N):
                for j in range(1, K - 1):
                    eps = max(abs(float(B[i, j] - A[i, j])), eps)
                    A[i, j] = B[i, j]
            eps0 = max(eps0, eps)
        lines.append('it=%d eps0=%f' % (it, eps0))
    return ''.join(x + '\n' for x in lines)


def test_tsk_lp_old_directive_form(tmp_path):
    """examples/tsk_lp.cdv (task regions in the old DVM(...) form, left to the
    preprocessor as cdvmh does): every line it prints equals the numpy
    restatement of the program."""
    out, _ = build_and_run('tsk_lp', tmp_path)
    assert out == tsk_lp_reference()
