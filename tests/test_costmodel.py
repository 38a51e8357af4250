"""Learned cost model: the HIP forward, backward and Adam kernels against the
float64 oracle (oracle/mlp_oracle.py), and what the trained network gains on a
loop site it was not trained on.

Tolerances: the kernels compute in FP32 with a different summation order than
the float64 oracle; gradients must agree within 1e-4 of the largest gradient
component (plus 1e-4 relative per component), outputs within 1e-4 relative."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle'))
import mlp_oracle as M  # noqa: E402

import dvmh_amd as D  # noqa: E402
from dvmh_amd import costmodel as CM  # noqa: E402


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def test_layout_constants_agree():
    """The oracle's parameter layout is the runtime's (count from the library)."""
    assert CM.rt().dvmh_amd_model_param_count_C() == M.NPARAM


def _case(n, seed):
    rng = np.random.default_rng(seed)
    p = f32(M.init_params(rng))
    X = f32(rng.normal(size=(n, M.F)))
    y = f32(rng.normal(size=n))
    return p, X, y


@pytest.mark.gpu
@pytest.mark.parametrize('n,seed', [(1, 1), (3, 2), (64, 3), (1000, 4), (20011, 5)])
def test_forward_matches_oracle(n, seed):
    p, X, y = _case(n, seed)
    out = np.zeros(n, dtype=np.float32)
    CM.rt().dvmh_amd_model_forward_C(CM.fptr(p), CM.fptr(X), n, CM.fptr(out))
    ref = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-4 * np.max(np.abs(ref)))


@pytest.mark.gpu
@pytest.mark.parametrize('n,seed', [(1, 1), (5, 2), (257, 3), (4096, 4), (20011, 5)])
def test_gradient_matches_oracle(n, seed):
    p, X, y = _case(n, seed)
    g = np.zeros(M.NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    CM.rt().dvmh_amd_model_grad_C(CM.fptr(p), CM.fptr(X), CM.fptr(y), n, CM.fptr(g), CM.fptr(loss))
    gref, lref = M.grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    scale = np.max(np.abs(gref))
    assert np.all(np.abs(g - gref) <= 1e-4 * scale + 1e-4 * np.abs(gref)), float(np.max(np.abs(g - gref)) / scale)
    assert abs(loss[0] - lref) <= 1e-4 * lref
    # deterministic: the same call gives the same bits
    g2 = np.zeros_like(g)
    CM.rt().dvmh_amd_model_grad_C(CM.fptr(p), CM.fptr(X), CM.fptr(y), n, CM.fptr(g2), None)
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))


@pytest.mark.gpu
def test_adam_matches_oracle():
    rng = np.random.default_rng(9)
    p = f32(rng.normal(size=M.NPARAM))
    m = np.zeros(M.NPARAM, dtype=np.float32)
    v = np.zeros(M.NPARAM, dtype=np.float32)
    pr, mr, vr = p.astype(np.float64), np.zeros(M.NPARAM), np.zeros(M.NPARAM)
    for step in range(1, 6):
        g = f32(rng.normal(size=M.NPARAM))
        CM.rt().dvmh_amd_model_adam_C(CM.fptr(p), CM.fptr(g), CM.fptr(m), CM.fptr(v), step, 1e-2, 0.9, 0.999, 1e-8)
        # the kernel's betas are FP32 values
        pr, mr, vr = M.adam_step(pr, g.astype(np.float64), mr, vr, step, 1e-2, beta1=float(np.float32(0.9)),
                                 beta2=float(np.float32(0.999)), eps=1e-8)
        assert np.allclose(p, pr, rtol=1e-5, atol=1e-6)
        assert np.allclose(m, mr, rtol=1e-5, atol=1e-7)
        assert np.allclose(v, vr, rtol=1e-5, atol=1e-9)


@pytest.mark.gpu
def test_training_fits_a_known_function():
    """Training through the runtime's kernels on synthetic records: a network
    trained by gradient steps of dvmh_amd_model_grad_C + adam_C (the kernels
    the runtime's trainer launches) fits y = sum of a few features."""
    rng = np.random.default_rng(21)
    n = 2048
    X = f32(rng.normal(size=(n, M.F)))
    y = f32(0.5 * X[:, 0] - 0.25 * X[:, 3] + 0.1 * X[:, 7] ** 2)
    p = f32(M.init_params(rng))
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    g = np.zeros_like(p)
    loss = np.zeros(1, dtype=np.float32)
    r = CM.rt()
    first = None
    for step in range(1, 301):
        r.dvmh_amd_model_grad_C(CM.fptr(p), CM.fptr(X), CM.fptr(y), n, CM.fptr(g), CM.fptr(loss))
        first = first if first is not None else float(loss[0])
        r.dvmh_amd_model_adam_C(CM.fptr(p), CM.fptr(g), CM.fptr(m), CM.fptr(v), step, 3e-3, 0.9, 0.999, 1e-8)
    final = M.loss(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert final < 0.05 * first, (first, final)


needs_records = pytest.mark.skipif(not os.path.exists(CM.RECORDS), reason='no recorded tuning sweep (costmodel/records.tsv)')


@pytest.mark.gpu
@needs_records
def test_model_gain_on_held_out_site():
    """Trained on the recorded sweep without the 384^3 Jacobi sites, the model's
    choice at a 384^3 Jacobi site (best measured of the default, its top 4 and
    its best candidate of any kernel variant those miss: a handful of timed
    executions instead of every candidate) is within 5% of the exhaustive
    optimum, beats the default configuration, and beats the same measuring
    budget spent on random candidates (the method without the model).  A
    training run is one draw (seeds 1-5 on these records land between 1.000
    and 1.057 of the optimum): the model of median quality over three seeds
    is the one checked."""
    evals = {}
    for seed in (1, 2, 3):
        CM.load_records(CM.RECORDS)
        n, loss = CM.train(epochs=3000, lr=3e-3, seed=seed, holdout=CM.HOLDOUT)
        assert n > 500 and loss < 0.5
        for site in [s for s in CM.sites() if CM.HOLDOUT in s and s.startswith('jac3d.cdv:')]:
            evals.setdefault(site, []).append(CM.evaluate(site, k=4))
    assert evals
    for site, evs in evals.items():
        ev = sorted(evs, key=lambda e: e['model_vs_exhaustive'])[1]
        print(site, [round(e['model_vs_exhaustive'], 4) for e in evs])
        assert ev['model_vs_exhaustive'] <= 1.05, ev
        assert ev['model_vs_exhaustive'] < ev['random_k_vs_exhaustive'], ev
        assert ev['model_speedup_vs_default'] > 1.0, ev


@pytest.mark.gpu
@needs_records
def test_trial_set_reaches_rare_variant():
    """The ADI sweeps along the innermost axis at 256^3 have kernel variants
    (the LDS-chunk and the streamed line sweeps) that few records show, so the
    model ranks them low; the trial set still times the best-ranked candidate
    of each, and the site's choice is
    within 10% of its exhaustive optimum (in-sample: these sites are trained on)."""
    CM.load_records(CM.RECORDS)
    CM.train(epochs=3000, lr=3e-3, seed=1, holdout=CM.HOLDOUT)
    sites = [s for s in CM.sites() if s.startswith(('adi3d.cdv:54|', 'adi3d.cdv:59|')) and s.endswith('e256x256x255')]
    assert sites
    for site in sites:
        ev = CM.evaluate(site, k=4)
        print(ev)
        assert ev['model_vs_exhaustive'] <= 1.10, ev
