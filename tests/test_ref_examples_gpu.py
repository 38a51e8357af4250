"""The reference's own C-DVMH example programs on the GPU: the executables
oracle/build_ref_examples.py built from examples/*.cdv (converted by the front
end, linked against libdvmh_amd) run with the GPU visible, so their regions
run on it, and what they print must equal what the host-path tests of
test_ref_examples.py require: the numpy restatements (bit for bit where the
program writes its grid), the golden outputs of examples/democ_res.uni, and
the benchmark programs' own verification lines (adi.cdv at 384^3, FP32
jac3d_perf.cdv at 384^3 and jac2d_perf.cdv at 8000^2, 100 iterations each, as
written).  A test is skipped when its executable was not built (the
reference tree is needed to build them)."""
import os

import numpy as np
import pytest

import dvmh_amd as D
import test_ref_examples as T

EX = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle', '_ref', 'examples')

pytestmark = pytest.mark.gpu


def run(name, tmp_path):
    exe = os.path.join(EX, name)
    if not os.path.exists(exe):
        pytest.skip(f'{name} was not built (oracle/build_ref_examples.py needs the reference tree)')
    env_req = os.environ.get('DVMH_AMD_REQUIRE_GPU')
    os.environ['DVMH_AMD_REQUIRE_GPU'] = '1'   # a region falling back to the host is an error
    try:
        return T.run_example(exe, str(tmp_path))
    finally:
        if env_req is None:
            del os.environ['DVMH_AMD_REQUIRE_GPU']
        else:
            os.environ['DVMH_AMD_REQUIRE_GPU'] = env_req


def test_gpu_visible():
    assert D.stats()['gpu_present'] == 1


@pytest.mark.parametrize('name,ref', [('jac2d', T.ref_jac2d), ('jacross', T.ref_jacross), ('jac3d', T.ref_jac3d),
                                      ('sor', T.ref_sor), ('redb_c', T.ref_redb)])
def test_example_program(name, ref, tmp_path):
    out, data = run(name, tmp_path)
    hist, final = ref()
    assert out == T.eps_lines(hist)
    if final is not None:
        assert data.shape == final.shape and np.array_equal(data.view(np.uint64), final.view(np.uint64))


def test_testcdvm(tmp_path):
    out, _ = run('testcdvm', tmp_path)
    assert out == 'T_CDVM  -  complete\n'


@pytest.mark.parametrize('name', ['gauss_c', 'gauss_wb'])
def test_gauss(name, tmp_path):
    out, _ = run(name, tmp_path)
    assert out == ''.join('j=%4i   X[j]=%e\n' % (j, 1.0) for j in range(8, -1, -1))


def test_mgrid(tmp_path):
    out, _ = run('mgrid', tmp_path)
    assert out == 'N1=8 M1=12 K1=16 \nN1=4 M1=6 K1=8 \n'


def test_tsk_lp(tmp_path):
    out, _ = run('tsk_lp', tmp_path)
    assert out == T.tsk_lp_reference()


def test_adi_20(tmp_path):
    out, _ = run('adi20', tmp_path)
    lines, eps = T.ref_adi(20)
    assert out.startswith(lines)


@pytest.mark.parametrize('name,last', [('adi', ' IT =  100   EPS =  7.2490737E-02\n'),
                                       ('jac3d_perf', ' IT =  100   EPS =  5.0580444E+00\n'),
                                       ('jac2d_perf', ' IT =  100   EPS =  5.8375977E+01\n')])
def test_benchmark_verification(name, last, tmp_path):
    out, _ = run(name, tmp_path)
    assert last in out
    assert ' Verification    =         SUCCESSFUL\n' in out
