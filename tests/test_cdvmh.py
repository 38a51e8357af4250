"""The C-DVMH front end (emit/cdvmh.py).

- Its loop descriptions of the configuration programs equal the hand-written
  ones (programs/*.loops, written from cdvmh's handler shape): loop variables,
  steps, parameters, reductions, ACROSS widths, shadow_renew arrays, bodies.
- The programs it converts from the .cdv sources (lib/libcdv_<prog>.so, the
  source's own functions) give the oracle's results: on the host path here,
  on the GPU in the `gpu` tests (bit-exact grids and MAX/MIN/MAXLOC histories,
  SUM within 1e-12 relative).
- relax2d.cdv (ACROSS on two axes, private, MAX + MIN, cuda_block, get_actual
  of a sub-array, host element reads) against its sequential semantics, the
  lexicographic order an ACROSS loop preserves.
- redall.cdv (SUM, PRODUCT, MAX, MIN, AND, OR, XOR, MAXLOC, MINLOC over
  double, long, int, unsigned int and float, initial values folded in, ties)
  against numpy folds of the same data (first occurrence for the LOC forms).
- Constructs outside the subset stop the conversion with the line.
"""
import ast
import math
import ctypes
import os
import re
import sys

import numpy as np
import pytest

import dvmh_amd as D
import helpers as H

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'dvm-snapshot-r6603-fork_amd')
sys.path.insert(0, os.path.join(PKG, 'emit'))
import cdvmh  # noqa: E402

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def _norm(body):
    return re.sub(r'\s+', ' ', body).strip()


# ---------------------------------------------------------------- descriptions
@pytest.mark.parametrize('prog', ['jac2d', 'jac3d', 'sor3d', 'adi3d'])
def test_descriptions_equal_hand_written(prog, tmp_path):
    cdvmh.convert_file(os.path.join(PKG, 'programs', prog + '.cdv'), str(tmp_path), prog=prog)
    got = ast.literal_eval(open(tmp_path / f'{prog}.loops').read().split('\n', 1)[1])
    want = ast.literal_eval(open(os.path.join(PKG, 'programs', prog + '.loops')).read())
    assert got['include'] == want['include']
    # loops in program order (the hand-written names of jac2d carry the pragma
    # line numbers of an earlier revision of the source)
    assert len(got['loops']) == len(want['loops'])
    for g, w in zip(got['loops'], want['loops']):
        assert g['vars'] == w['vars']
        assert g['steps'] == w.get('steps', ['1'] * len(w['vars']))
        assert sorted(g['params']) == sorted(w['params'])
        assert g.get('reductions', []) == w.get('reductions', [])
        assert g.get('across') == w.get('across')
        assert sorted(g.get('stencil', [])) == sorted(w.get('stencil', []))
        assert _norm(g['body']) == _norm(w['body'])


def test_converted_text_calls(tmp_path):
    """The converted relax2d: declaration, allocation, distribution, region
    registrations with inferred intents, loop registration, sub-array
    get_actual, element reads, release."""
    cdvmh.convert_file(os.path.join(PKG, 'programs', 'relax2d.cdv'), str(tmp_path), prog='relax2d_cdv')
    c = open(tmp_path / 'relax2d_cdv.c').read()
    for call in ['dvmh_array_declare_C(A, 2, -rt_DOUBLE, (DvmType)(0), (DvmType)1, (DvmType)1, (DvmType)(n)',
                 'dvmh_array_alloc_C(A, n * n * sizeof(double));',
                 'dvmh_distribute_C(A, 2, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2));',
                 'dvmh_region_register_array_C(dvmh_region_1, INTENT_OUT, A, "A");',
                 'dvmh_region_register_array_C(dvmh_region_2, INTENT_INOUT, A, "A");',
                 'dvmh_region_register_scalar_C(dvmh_region_2, INTENT_IN, &w, sizeof(w), "w");',
                 'dvmh_loop_set_cuda_block_C(dvmh_loop_1, 64, 2, 1);',
                 'dvmh_loop_across_C(dvmh_loop_2, 0, A, 2, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1);',
                 'dvmh_loop_reduction_C(dvmh_loop_2, rf_MIN, &mn, rt_DOUBLE, 1, 0, 0);',
                 'dvmh_get_actual_subarray_C(A, 2, (DvmType)(1), (DvmType)(n - 2), (DvmType)(0), (DvmType)(n - 1));',
                 '(*(double *)dvmh_get_element_addr_C(A, 2, (DvmType)(i), (DvmType)(j)))',
                 'dvmh_array_free_C(A);', 'dvmh_forget_header_(A);']:
        assert call in c, call
    spec = ast.literal_eval(open(tmp_path / 'relax2d_cdv.loops').read().split('\n', 1)[1])
    sweep = spec['loops'][1]
    assert sweep['privates'] == ['double t'] and sweep['across']['widths'] == [(1, 1), (1, 1)]
    assert [p[1] for p in sweep['params']] == ['A', 'w']


BAD = [
    ('nested regions', '#pragma dvm region\n{\n#pragma dvm region\n{\n}\n}\n', 'nested regions'),
    ('imperfect nest', '#pragma dvm region\n{\n#pragma dvm parallel([i][j] on A[i][j])\nfor (i = 0; i < n; i++) {\n'
     'x = 1;\nfor (j = 0; j < n; j++) A[i][j] = 0;\n}\n}\n', 'perfect'),
    ('unknown clause', '#pragma dvm region\n{\n#pragma dvm parallel([i][j] on A[i][j]) shadow_compute\n'
     'for (i = 0; i < n; i++) for (j = 0; j < n; j++) A[i][j] = 0;\n}\n', 'shadow_compute'),
    ('non-linear subscript', '#pragma dvm region\n{\n#pragma dvm parallel([i][j] on A[i * j][j])\n'
     'for (i = 0; i < n; i++) for (j = 0; j < n; j++) A[i][j] = 0;\n}\n', 'subscript'),
    ('C array in a region', '#pragma dvm region\n{\n#pragma dvm parallel([i][j] on A[i][j])\n'
     'for (i = 0; i < n; i++) for (j = 0; j < n; j++) A[i][j] = C[i][j];\n}\n', 'inside regions'),
]


@pytest.mark.parametrize('what,body,msg', BAD, ids=[b[0] for b in BAD])
def test_unsupported_constructs_name_the_line(what, body, msg):
    src = ('void f(long n)\n{\n    int i, j, x;\n    double C[8][8];\n    #pragma dvm array distribute[block][block]\n    double (*A)[n];\n'
           '    A = malloc(n * n * sizeof(double));\n' + body + '}\n')
    with pytest.raises(cdvmh.ConvertError) as e:
        cv = cdvmh.Converter(src, 't.cdv', 't')
        cv.convert()
    assert 't.cdv:' in str(e.value) and msg in str(e.value)


# ---------------------------------------------------------------- converted programs
def _lib(prog):
    return D.load(f'libcdv_{prog}.so')


def run_cdv(prog, n, iters):
    lib = _lib(prog)
    if prog == 'jac2d':
        f = lib.jac2d
        f.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_ulonglong, _dp]
        eps = np.zeros(iters)
        f(n, iters, D.SEED_BASE + 1, H.dp(eps))
        return {'eps': eps}
    if prog == 'jac3d':
        f = lib.jac3d
        f.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_ulonglong, _dp, _dp]
        eps, s = np.zeros(iters), np.zeros(iters)
        f(n, iters, D.SEED_BASE + 2, H.dp(eps), H.dp(s))
        return {'eps': eps, 'sum': s}
    if prog == 'sor3d':
        f = lib.sor3d
        f.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp]
        eps = np.zeros(iters)
        f(n, iters, 1.5, D.SEED_BASE + 3, H.dp(eps))
        return {'eps': eps}
    f = lib.adi3d
    f.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp, _lp]
    mx, loc = np.zeros(iters), np.zeros(iters, dtype=np.int64)
    f(n, iters, 0.3, D.SEED_BASE + 4, H.dp(mx), loc.ctypes.data_as(_lp))
    return {'max': mx, 'loc': loc}


def check_against_oracle(prog, n, iters):
    got = run_cdv(prog, n, iters)
    if prog == 'jac2d':
        assert np.array_equal(got['eps'], H.oracle_jac2d(n, iters)['eps'])
    elif prog == 'jac3d':
        ref = H.oracle_jac3d(n, iters)
        assert np.array_equal(got['eps'], ref['eps']) and H.sum_close(got['sum'], ref['sum'])
    elif prog == 'sor3d':
        assert np.array_equal(got['eps'], H.oracle_sor3d(n, iters)['eps'])
    else:
        ref = H.oracle_adi3d(n, iters)
        assert np.array_equal(got['max'], ref['max']) and np.array_equal(got['loc'], ref['loc'])


HOST_CASES = [('jac2d', 20, 3), ('jac3d', 12, 3), ('sor3d', 9, 2), ('adi3d', 8, 2)]


@pytest.mark.parametrize('prog,n,iters', HOST_CASES)
def test_converted_program_host_path(prog, n, iters):
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    check_against_oracle(prog, n, iters)


def _gpu_loops():
    return D.stats()['loops_gpu']


GPU_CASES = [('jac2d', 5, 3), ('jac2d', 70, 4), ('jac2d', 1000, 3), ('jac3d', 10, 3), ('jac3d', 66, 4),
             ('jac3d', 200, 3), ('sor3d', 5, 2), ('sor3d', 23, 3), ('sor3d', 70, 2), ('adi3d', 6, 2), ('adi3d', 33, 2),
             ('adi3d', 64, 3)]


@pytest.mark.gpu
@pytest.mark.parametrize('prog,n,iters', GPU_CASES)
def test_converted_program_gpu(prog, n, iters):
    before = _gpu_loops()
    check_against_oracle(prog, n, iters)
    assert _gpu_loops() > before


# ---------------------------------------------------------------- relax2d
def relax2d_reference(n, iters, w, seed):
    """Sequential execution of relax2d.cdv: lexicographic Gauss-Seidel, the
    order the ACROSS loop preserves."""
    A = np.zeros(n * n)
    H.oracle().orc_init_grid2d(H.vp(A), n, seed, 1.0)
    a = A.reshape(n, n).tolist()
    res_h, min_h = [], []
    for _ in range(iters):
        res, mn = 0.0, 1e300
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                t = 0.25 * (a[i - 1][j] + a[i + 1][j] + a[i][j - 1] + a[i][j + 1])
                t = a[i][j] + w * (t - a[i][j])
                d = abs(t - a[i][j])
                res = res if res > d else d
                mn = mn if mn < t else t
                a[i][j] = t
        res_h.append(res)
        min_h.append(mn)
    return np.array(res_h), np.array(min_h), np.array(a[1:n - 1]).ravel()


def run_relax2d(n, iters, w, seed):
    f = _lib('relax2d').relax2d
    f.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _dp, _dp, _dp]
    res, mn, out = np.zeros(iters), np.zeros(iters), np.zeros((n - 2) * n)
    f(n, iters, w, seed, H.dp(res), H.dp(mn), H.dp(out))
    return res, mn, out


def _check_relax2d(n, iters, w=1.25, seed=D.SEED_BASE + 7):
    res, mn, out = run_relax2d(n, iters, w, seed)
    rr, rm, ro = relax2d_reference(n, iters, w, seed)
    assert np.array_equal(res, rr) and np.array_equal(mn, rm)
    assert H.bits_equal(out, ro)


@pytest.mark.parametrize('n,iters', [(3, 2), (16, 3)])
def test_relax2d_host_path(n, iters):
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    _check_relax2d(n, iters)


@pytest.mark.gpu
@pytest.mark.parametrize('n,iters', [(3, 2), (4, 2), (17, 3), (65, 2), (130, 2)])
def test_relax2d_gpu(n, iters):
    """ACROSS on both axes of a 2-D loop: one hyperplane launch per
    anti-diagonal, MAX and MIN in one loop, and the sub-array fetch."""
    before = _gpu_loops()
    _check_relax2d(n, iters)
    assert _gpu_loops() > before


# ---------------------------------------------------------------- redall
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    """SplitMix64 finaliser (programs/dvmh_prog_common.h prog_mix64) on uint64 arrays."""
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _u01(seed, idx):
    """prog_u01 of programs/dvmh_prog_common.h on uint64 index arrays."""
    z = np.uint64(seed) + (idx + np.uint64(1)) * _GOLD
    return (_mix64(z) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def test_generator_port_matches_oracle_grid():
    """The numpy port of the programs' generator equals the oracle's (interior
    points of orc_init_grid2d are prog_u01(seed, i * n + j))."""
    n, seed = 9, D.SEED_BASE + 8
    g = np.zeros(n * n)
    H.oracle().orc_init_grid2d(H.vp(g), n, seed, 1.0)
    u = _u01(seed, np.arange(n * n, dtype=np.uint64)).reshape(n, n)
    assert H.bits_equal(g.reshape(n, n)[1:-1, 1:-1], u[1:-1, 1:-1])


def redall_reference(n, m, seed):
    """Results of redall.cdv from numpy on the same data: library folds, first
    occurrence for MAXLOC/MINLOC (numpy.argmax/argmin), each reduction's
    initial value folded in as the program's comment states."""
    idx = np.arange(n * m, dtype=np.uint64)
    z = _mix64(np.uint64(seed) + idx)
    A = _u01(seed, idx)
    K = (z % np.uint64(1000)).astype(np.int64) - 500
    I = ((z >> np.uint64(20)) % np.uint64(300)).astype(np.int64)
    U = z >> np.uint64(32)
    F = _u01(seed + 1, idx).astype(np.float32)
    w = (_mix64(np.uint64(seed + 7) + idx) | np.uint64(0x0101000000000000)).view(np.int64)
    return {
        'dsum': math.fsum(list(A) + [1.5]),
        'dprod': 2.0 * math.exp(math.fsum(np.log1p((A - 0.5) * 0.001))),
        'dmax': max(float(A.max()), 0.25), 'dmn': float(A.min()), 'dloc': int(A.argmin()),
        'lsum': int(K.sum()) + 1000000, 'lmax': max(int(K.max()), 10000), 'lmin': min(int(K.min()), -450),
        'land': int(np.bitwise_and.reduce(w)) & ~0xFF, 'lor': int(np.bitwise_or.reduce(w & 0xFF00)) | 0x10,
        'lxor': int(np.bitwise_xor.reduce(w)) ^ 0x5A5A, 'kmx': int(K.max()), 'kloc': int(K.argmax()),
        'isum': int(I.sum()) + 7, 'imn': int(I.min()), 'iloc': int(I.argmin()),
        'umax': max(int(U.max()), 5), 'umin': min(int(U.min()), 0xFFFFFFFF),
        'fsum': math.fsum(F.astype(np.float64)) + 0.5, 'fmax': float(F.max()),
        'ties': (int((K == K.max()).sum()), int((I == I.min()).sum())),
    }


def run_redall(n, m, reps, seed):
    f = _lib('redall').redall
    f.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_ulonglong, _dp, _lp,
                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_float)]
    dres, lres = np.zeros(4), np.zeros(9, dtype=np.int64)
    ires, ures, fres = np.zeros(3, dtype=np.int32), np.zeros(2, dtype=np.uint32), np.zeros(2, dtype=np.float32)
    f(n, m, reps, seed, H.dp(dres), lres.ctypes.data_as(_lp), ires.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
      ures.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), fres.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    names = ['dsum', 'dprod', 'dmax', 'dmn'], ['dloc', 'lsum', 'lmax', 'lmin', 'land', 'lor', 'lxor', 'kmx', 'kloc'], \
        ['isum', 'imn', 'iloc'], ['umax', 'umin'], ['fsum', 'fmax']
    out = {}
    for ns, arr in zip(names, (dres, lres, ires, ures, fres)):
        for k, v in zip(ns, arr):
            out[k] = v.item()
    return out


def _check_redall(n, m, reps=2, seed=D.SEED_BASE + 8):
    got, ref = run_redall(n, m, reps, seed), redall_reference(n, m, seed)
    exact = [k for k in got if k not in ('dsum', 'dprod', 'fsum')]
    assert {k: got[k] for k in exact} == {k: ref[k] for k in exact}
    # SUM and PRODUCT in FP64 within 1e-12 relative of the exact value; SUM in
    # FP32 within 1e-4 relative (n * m float additions in the loop's order)
    assert abs(got['dsum'] - ref['dsum']) <= 1e-12 * abs(ref['dsum'])
    assert abs(got['dprod'] - ref['dprod']) <= 1e-12 * abs(ref['dprod'])
    assert abs(got['fsum'] - ref['fsum']) <= 1e-4 * abs(ref['fsum'])
    return ref


@pytest.mark.parametrize('n,m', [(1, 1), (3, 5), (37, 61), (256, 300)])
def test_redall_host_path(n, m):
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    ref = _check_redall(n, m)
    if n * m > 1000:
        assert min(ref['ties']) > 1   # MAXLOC / MINLOC ties are exercised


@pytest.mark.gpu
@pytest.mark.parametrize('n,m', [(1, 1), (3, 5), (37, 61), (256, 300), (1000, 1037), (2048, 2500)])
def test_redall_gpu(n, m):
    """Every reduction function over double, long, int, unsigned and float on
    the GPU path: one-block and multi-block finishes, element types grouped
    per finish launch, MAXLOC/MINLOC ties across workgroups."""
    before = _gpu_loops()
    ref = _check_redall(n, m)
    if n * m > 1000:
        assert min(ref['ties']) > 1
    assert _gpu_loops() > before


# ---------------------------------------------------------------- jacg
def _check_jacg(iters, seed=D.SEED_BASE + 1):
    """jacg.cdv (file-scope arrays, loop variables and macros, shadow widths,
    whole arrays handed to host code) against the oracle's 2-D Jacobi."""
    n = 131
    f = _lib('jacg').jacg
    f.argtypes = [ctypes.c_int, ctypes.c_ulonglong, _dp, _dp, _dp]
    eps, A, B = np.zeros(iters), np.zeros(n * n), np.zeros(n * n)
    f(iters, seed, H.dp(eps), H.dp(A), H.dp(B))
    ref = H.oracle_jac2d(n, iters, seed)
    assert np.array_equal(eps, ref['eps'])
    assert H.bits_equal(A, ref['A']) and H.bits_equal(B, ref['B'])


def test_jacg_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    _check_jacg(4)


@pytest.mark.gpu
def test_jacg_gpu():
    before = _gpu_loops()
    _check_jacg(6)
    assert _gpu_loops() > before


def test_host_loop_with_c_array():
    """A parallel loop outside regions runs on the host only: an ordinary C
    array of the program reaches its host handler by address, and no GPU
    handler is made for it.  (A file-scope initialiser in braces is passed
    through.)"""
    src = ('#define M 6\nstatic const int W[3] = {1, 2, 3};\nvoid f(long n)\n{\n    int i, j, s;\n    int C[5][M];\n'
           '    #pragma dvm array distribute[block][block]\n    int (*A)[M];\n    A = malloc(n * M * sizeof(int));\n'
           '    s = 0;\n    #pragma dvm parallel([i][j] on A[i][j]) reduction(sum(s))\n'
           '    for (i = 0; i < 5; i++)\n        for (j = 0; j < M; j++)\n            s = s + A[i][j] * C[i][j];\n}\n')
    cv = cdvmh.Converter(src, 't.cdv', 't')
    text = cv.program_text(cv.convert())
    (L,) = cv.loops
    assert L['host_only'] and ('carray', 'C', 'int', ['5', 'M']) in L['params']
    assert 'void loop_t_11_host(DvmType *, DvmType[], void *);' in text
    assert 'loop_t_11_cuda' not in text
    assert 'dvmh_handler_func_C((DvmHandlerFunc)loop_t_11_host, 2, A, C)' in text
    sys.path.insert(0, os.path.join(PKG, 'emit'))
    import hipgen
    host = hipgen.host_handler(hipgen.Loop(L, 't'))
    assert 'int (*C)[M] = (int (*)[M])C_ptr;' in host


def test_file_scope_conversion(tmp_path):
    """File-scope arrays become headers created once by the first converted
    function that runs; shadow widths reach the creation call; a whole array
    used as a value is its host copy; an inherited parameter is the caller's
    header."""
    cdvmh.convert_file(os.path.join(PKG, 'programs', 'jacg.cdv'), str(tmp_path), prog='jacg_cdv')
    c = open(tmp_path / 'jacg_cdv.c').read()
    for call in ['DvmType A[64];', 'DvmType B[64];', 'static void jacg_cdv_globals_init(void);',
                 'dvmh_array_create_C(A, 2, -rt_DOUBLE, (DvmType)(N), (DvmType)1, (DvmType)1, (DvmType)(N), (DvmType)1, (DvmType)1);',
                 'copy_out(a_out, ((void *)dvmh_get_natural_base_C(0, A)));',
                 # inherited parameters: headers in both declarations, passed as headers
                 'static double sweep(DvmType P[64], DvmType Q[64]);',
                 'static double sweep(DvmType P[64], DvmType Q[64])\n{',
                 'eps_hist[it - 1] = sweep(A, B);',
                 'dvmh_region_register_array_C(dvmh_region_2, INTENT_INOUT, P, "P");']:
        assert call in c, call
    spec = ast.literal_eval(open(tmp_path / 'jacg_cdv.loops').read().split('\n', 1)[1])
    assert '#define MAXOF(a, b) ((a) > (b) ? (a) : (b))' in spec['defines']


# ---------------------------------------------------------------- wide: 256 handler parameters
def _check_wide(n, m):
    """programs/wide.cdv: one loop whose handlers receive 256 parameters (the
    array, m and 254 weighted scalars); the result against a Python evaluation
    of the same expression in the same order."""
    rng = np.random.default_rng(n * 1000 + m)
    c = rng.uniform(-1.0, 1.0, 254)
    acc = c[0] * 1.0
    for k in range(1, len(c)):
        acc = acc + c[k] * float(k + 1)
    want = np.array([acc + float(q) for q in range(n * m)])
    f = _lib('wide').wide
    f.argtypes = [ctypes.c_long, ctypes.c_long, _dp, _dp]
    out = np.zeros(n * m)
    f(n, m, H.dp(c), H.dp(out))
    assert H.bits_equal(out, want)


def test_wide_handler_host_path():
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    _check_wide(5, 7)


@pytest.mark.gpu
@pytest.mark.parametrize('n,m', [(1, 1), (33, 65)])
def test_wide_handler_gpu(n, m):
    before = _gpu_loops()
    _check_wide(n, m)
    assert _gpu_loops() > before


# ---------------------------------------------------------------- redarr: reduction arrays
def redarr_reference(n, seed):
    """Sequential execution of programs/redarr.cdv (numpy for the PARALLEL
    loop, loops in program order for the ACROSS ones)."""
    A = _u01(seed, np.arange(n * n, dtype=np.uint64)).reshape(n, n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    s = np.bincount(((i + j) % 5).ravel(), weights=A.ravel(), minlength=5)
    mx = np.full(3, 0.75)
    for c in range(3):
        mx[c] = max(mx[c], np.max(A[:, c::3] * float(c + 1), initial=-np.inf))
    mn = np.full(4, 0.25)
    for r in range(4):
        mn[r] = min(mn[r], np.min(A[r::4, :], initial=np.inf))
    cnt = np.array([0, 10, 20], dtype=np.int64) + np.bincount(((i * n + j) % 3).ravel(), minlength=3)
    lm = np.array([-1.0, 2.5])
    for ii in range(1, n):
        A[ii, :] = A[ii, :] + 0.5 * A[ii - 1, :]
        for c in range(2):
            lm[c] = max(lm[c], np.max(A[ii, c::2], initial=-np.inf))
    wm = np.zeros(3)
    for ii in range(1, n):
        for jj in range(1, n):
            A[ii, jj] = 0.5 * (A[ii - 1, jj] + A[ii, jj - 1]) + 0.25 * A[ii, jj]
            q = (ii + 2 * jj) % 3
            wm[q] = max(wm[q], A[ii, jj])
    return {'sums': s, 'maxs': mx, 'mins': mn, 'counts': cnt, 'lmax': lm, 'wmax': wm, 'A': A.ravel()}


def _check_redarr(n):
    seed = D.SEED_BASE + 17
    f = _lib('redarr').redarr
    lp = ctypes.POINTER(ctypes.c_long)
    f.argtypes = [ctypes.c_long, ctypes.c_ulonglong, _dp, _dp, _dp, lp, _dp, _dp, _dp]
    got = {'sums': np.zeros(5), 'maxs': np.zeros(3), 'mins': np.zeros(4), 'counts': np.zeros(3, dtype=np.int64),
           'lmax': np.zeros(2), 'wmax': np.zeros(3), 'A': np.zeros(n * n)}
    f(n, seed, H.dp(got['sums']), H.dp(got['maxs']), H.dp(got['mins']), got['counts'].ctypes.data_as(lp),
      H.dp(got['lmax']), H.dp(got['wmax']), H.dp(got['A']))
    want = redarr_reference(n, seed)
    assert H.bits_equal(got['A'], want['A'])
    assert H.sum_close(got['sums'], want['sums'])
    for k in ('maxs', 'mins', 'counts', 'lmax', 'wmax'):
        assert np.array_equal(got[k], want[k]), (k, got[k], want[k])


@pytest.mark.parametrize('n', [2, 9])
def test_redarr_host_path(n):
    if D.stats()['gpu_present']:
        pytest.skip('the host path runs where no GPU is present; the GPU tests cover this case')
    _check_redarr(n)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [2, 3, 17, 70, 131])
def test_redarr_gpu(n):
    """Reduction arrays on the GPU: per-workgroup partials of every element
    (element-major), folded per element by the multi-reduction finish; ACROSS
    loops accumulate into pre-filled element slots."""
    before = _gpu_loops()
    _check_redarr(n)
    assert _gpu_loops() > before
