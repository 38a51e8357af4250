"""Learned cost model workflow around libdvmh_amd's tuner.

  sweep   run workloads with the exhaustive tuner (every candidate launch
          configuration of every loop site timed) and save the records;
  train   train the perceptron on the GPU from the records (the runtime's own
          HIP forward/backward/Adam kernels), optionally holding out sites;
  eval    for a held-out site, compare the model's choice (best measured of
          its top-k predictions plus the default) with the exhaustive optimum,
          the default configuration and k random candidates.

    python -m dvmh_amd.costmodel sweep --out records.tsv
    python -m dvmh_amd.costmodel sweep --merge --spec adi3d:64,128 --out records.tsv
    python -m dvmh_amd.costmodel train --records records.tsv --out costmodel.bin --holdout '|e382x382x382'
    python -m dvmh_amd.costmodel eval --records records.tsv --model costmodel.bin --site 'jac3d.cdv:41|k3|r3|m0|e382x382x382'
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import runtime, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(os.path.dirname(HERE), 'costmodel')
RECORDS = os.path.join(DATA_DIR, 'records.tsv')
MODEL = os.path.join(DATA_DIR, 'costmodel.bin')

# Sweep: several sizes of every workload kind (PARALLEL 2-D and 3-D stencils,
# LINE sweeps, WAVEFRONT).  The 384^3 Jacobi sites (extents 382^3) are recorded
# too but held out of training (HOLDOUT), so they measure generalisation.
DEFAULT_SPECS = [
    ('jac3d', [64, 96, 128, 160, 192, 256, 320, 384, 448]),
    ('jac2d', [256, 512, 1024, 1536, 2048, 3072, 4096, 6144]),
    ('adi3d', [64, 96, 128, 192, 256]),
    ('sor3d', [48, 64, 96, 128]),
]
HOLDOUT = '|e382x382x382'
HELD_OUT_SITE = 'jac3d.cdv:41|k3|r3|m0|e382x382x382'

_bound = None


def rt():
    global _bound
    if _bound is None:
        lib = runtime()
        L, Dd, cp = ctypes.c_long, ctypes.c_double, ctypes.c_char_p
        fp = ctypes.POINTER(ctypes.c_float)
        dpp = ctypes.POINTER(ctypes.c_double)
        sig = {
            'dvmh_amd_tuner_set_repeats_C': ([L], None),
            'dvmh_amd_set_graph_C': ([L], None),
            'dvmh_amd_tuner_clear_C': ([], None),
            'dvmh_amd_tuner_site_pending_C': ([cp, L], L),
            'dvmh_amd_tuner_sites_C': ([ctypes.c_char_p, L], L),
            'dvmh_amd_model_param_count_C': ([], L),
            'dvmh_amd_model_train_C': ([L, Dd, L, cp, dpp], L),
            'dvmh_amd_model_save_C': ([cp], L),
            'dvmh_amd_model_load_C': ([cp], L),
            'dvmh_amd_model_ready_C': ([], L),
            'dvmh_amd_model_site_predict_C': ([cp, dpp, dpp, L], L),
            'dvmh_amd_model_site_features_C': ([cp, fp, L], L),
            'dvmh_amd_model_grad_C': ([fp, fp, fp, L, fp, fp], L),
            'dvmh_amd_model_forward_C': ([fp, fp, L, fp], L),
            'dvmh_amd_model_adam_C': ([fp, fp, fp, fp, L, Dd, Dd, Dd, Dd], L),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _bound = lib
    return _bound


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sites():
    lib = rt()
    n = lib.dvmh_amd_tuner_sites_C(None, 0)
    buf = ctypes.create_string_buffer(int(n))
    lib.dvmh_amd_tuner_sites_C(buf, n)
    return [s for s in buf.value.decode().split('\n') if s]


def sweep_workload(name, size, repeats=2, iters_per_run=20, max_runs=400, log=None):
    """Runs one workload under the exhaustive tuner until every candidate of
    its iteration-phase loop sites has `repeats` timings."""
    lib = rt()
    lib.dvmh_amd_tuner_set_mode_C(1)
    lib.dvmh_amd_tuner_set_repeats_C(repeats)
    w = workloads.ALL[name](size)
    lib.dvmh_amd_set_graph_C(2 if w.lazy else 0)
    w.setup()
    runs = 0
    t0 = time.perf_counter()
    while True:
        w.run(iters_per_run)
        runs += 1
        pending = sum(lib.dvmh_amd_tuner_site_pending_C(w.cdv, ln) for ln in w.lines)
        if pending == 0:
            break
        if runs >= max_runs:
            raise RuntimeError(f'{name} {size}: tuner still has {pending} candidates pending after {runs} runs')
    w.teardown()
    lib.dvmh_amd_tuner_set_mode_C(0)
    lib.dvmh_amd_set_graph_C(0)
    if log:
        log(f'swept {name} {size}: {runs} runs, {time.perf_counter() - t0:.1f} s')


def sweep(specs=DEFAULT_SPECS, out=RECORDS, repeats=2, log=None, merge=False):
    """Sweeps the workloads from a clean tuner and saves the records to `out`.
    With `merge`, `out` keeps its records of every other program: those of the
    swept workloads' programs are replaced by the new sweep."""
    lib = rt()
    lib.dvmh_amd_tuner_clear_C()
    programs = set()
    for name, sizes in specs:
        for size in sizes:
            sweep_workload(name, size, repeats=repeats, log=log)
            W = workloads.ALL[name]
            programs.update(W.record_prefixes or (W.cdv.decode() + ':',))
    os.makedirs(os.path.dirname(out) or '.', exist_ok=True)
    kept = []
    if merge and os.path.exists(out):
        with open(out) as f:
            kept = [ln for ln in f if not any(ln.startswith(p) for p in programs)]
    tmp = out + '.new'
    n = lib.dvmh_amd_tuner_save_C(tmp.encode())
    with open(tmp) as f:
        fresh = f.readlines()
    with open(out, 'w') as f:
        f.writelines(kept + fresh)
    os.remove(tmp)
    if log and merge:
        log(f'kept {len(kept)} records of other programs')
    return len(kept) + n


def load_records(path=RECORDS, clear=True):
    lib = rt()
    if clear:
        lib.dvmh_amd_tuner_clear_C()
    n = lib.dvmh_amd_tuner_load_C(path.encode())
    if n < 0:
        raise FileNotFoundError(path)
    return n


def train(epochs=3000, lr=3e-3, seed=1, holdout=HOLDOUT):
    lib = rt()
    loss = ctypes.c_double(0)
    n = lib.dvmh_amd_model_train_C(epochs, lr, seed, holdout.encode() if holdout else None, ctypes.byref(loss))
    return n, loss.value


def site_table(site):
    """(predicted seconds or None, measured seconds) for every candidate of the site."""
    lib = rt()
    n = lib.dvmh_amd_model_site_predict_C(site.encode(), None, None, 0)
    if n < 0:
        raise KeyError(site)
    meas = np.zeros(n)
    lib.dvmh_amd_model_site_predict_C(site.encode(), None, dptr(meas), n)
    pred = None
    if lib.dvmh_amd_model_ready_C():
        pred = np.zeros(n)
        lib.dvmh_amd_model_site_predict_C(site.encode(), dptr(pred), None, n)
    return pred, meas


def expected_random_best(meas, k, trials=20000, seed=0):
    """Expected time of the best of the default and k distinct random other
    candidates (the tuner without a model, given the same measuring budget)."""
    rng = np.random.default_rng(seed)
    others = meas[1:]
    best = np.empty(trials)
    for t in range(trials):
        pick = rng.choice(len(others), size=min(k, len(others)), replace=False)
        best[t] = min(meas[0], others[pick].min())
    return float(best.mean())


VARIANT_FEATURE = 14   # feature column of the kernel variant (model.cpp features())


def site_variants(site, n):
    x = np.zeros((n, 24), dtype=np.float32)
    rt().dvmh_amd_model_site_features_C(site.encode(), fptr(x), n)
    return x[:, VARIANT_FEATURE].astype(int)


def trial_set(pred, variants, k=4):
    """The candidates the tuner times in cost-model mode (tuner.cpp
    SiteInfo::trialSet): the default, the model's best k, and the model's
    best candidate of every kernel variant those miss."""
    order = [int(c) for c in np.argsort(pred, kind='stable')]
    top = [0] + [c for c in order if c != 0][:k]
    for c in order:
        if variants[c] not in {variants[t] for t in top}:
            top.append(c)
    return top


def evaluate(site, k=4):
    """Quality of the model's ranking on a site with complete records."""
    pred, meas = site_table(site)
    if np.any(meas <= 0):
        raise ValueError(f'site {site} has unmeasured candidates')
    best = float(meas.min())
    top = trial_set(pred, site_variants(site, len(meas)), k)
    chosen = float(meas[top].min())
    rand = expected_random_best(meas, len(top) - 1)   # the same number of timed candidates
    rp = np.argsort(np.argsort(pred))
    rm = np.argsort(np.argsort(meas))
    spearman = float(np.corrcoef(rp, rm)[0, 1])
    return {'site': site, 'candidates': int(len(meas)), 'k': k, 'exhaustive_s': best, 'default_s': float(meas[0]),
            'model_s': chosen, 'model_vs_exhaustive': chosen / best, 'default_vs_exhaustive': float(meas[0]) / best,
            'random_k_vs_exhaustive': rand / best, 'model_speedup_vs_default': float(meas[0]) / chosen,
            'spearman': spearman, 'model_top': top, 'trials': len(top)}


def main(argv=None):
    ap = argparse.ArgumentParser(prog='python -m dvmh_amd.costmodel')
    sub = ap.add_subparsers(dest='cmd', required=True)
    s = sub.add_parser('sweep')
    s.add_argument('--out', default=RECORDS)
    s.add_argument('--repeats', type=int, default=2)
    s.add_argument('--spec', action='append', help='name:size,size,...  (default: the training sweep)')
    s.add_argument('--merge', action='store_true', help='replace only the swept programs\' records in --out')
    t = sub.add_parser('train')
    t.add_argument('--records', default=RECORDS)
    t.add_argument('--out', default=MODEL)
    t.add_argument('--holdout', default=HOLDOUT)
    t.add_argument('--epochs', type=int, default=3000)
    t.add_argument('--lr', type=float, default=3e-3)
    t.add_argument('--seed', type=int, default=1)
    e = sub.add_parser('eval')
    e.add_argument('--records', default=RECORDS)
    e.add_argument('--model', default=MODEL)
    e.add_argument('--site', default=HELD_OUT_SITE)
    e.add_argument('-k', type=int, default=4)
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, flush=True)

    if a.cmd == 'sweep':
        specs = DEFAULT_SPECS
        if a.spec:
            specs = [(x.split(':')[0], [int(v) for v in x.split(':')[1].split(',')]) for x in a.spec]
        n = sweep(specs, a.out, a.repeats, log, merge=a.merge)
        log(f'{n} records -> {a.out}')
    elif a.cmd == 'train':
        load_records(a.records)
        n, loss = train(a.epochs, a.lr, a.seed, a.holdout)
        if n == 0:
            sys.exit('no training samples')
        rt().dvmh_amd_model_save_C(a.out.encode())
        log(f'trained on {n} samples, loss {loss:.4f} -> {a.out}')
    else:
        load_records(a.records)
        if rt().dvmh_amd_model_load_C(a.model.encode()) < 0:
            sys.exit(f'cannot load {a.model}')
        log(evaluate(a.site, a.k))


if __name__ == '__main__':
    main()
