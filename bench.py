#!/usr/bin/env python3
"""Benchmark: 3-D 7-point Jacobi, FP64 384^3, 50 iterations per step, through
the DVMH C interface on one MI355X.

A step is one run of the program's iteration phase (50 iterations, each one
PARALLEL loop with REDUCTION(MAX) and REDUCTION(SUM)) on inputs already
resident in HBM.  Two forms run side by side:
  plain   the default launch configuration, one synchronised launch sequence
          per loop (what the reference runtime does);
  chosen  the tuner's configuration (learned cost model when a trained model
          is present, else the exhaustive sweep done during warm-up), results
          synchronised only when needed, repeating regions replayed as graphs.
`value` is the chosen form's point-updates/s.  The roofline uses the stencil
kernel's average duration measured with HIP events on the runtime's stream
during the timed region.  The CPU baseline is the oracle (the C/OpenMP
restatement of the host handlers) timed on a bounded sample of the same
workload on this machine's host cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, 'dvm-snapshot-r6603-fork_amd')
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(REPO, 'tests'))

import numpy as np  # noqa: E402

import dvmh_amd as D  # noqa: E402

METRIC = 'FP64 point-updates/s + % HBM roofline; cost-model/batched vs default, bit-exact'
N, ITERS = 384, 50
SEED = D.SEED_BASE + 2
HBM_PEAK = 8.0e12
B_ALG = 16          # bytes per point-update: read A, write B (8 B each)
PTS = (N - 2) ** 3  # point-updates per iteration
STENCIL_LINES = (41, 54)
MODEL_FILE = os.path.join(PKG, 'models', 'costmodel.bin')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Jac3d:
    def __init__(self):
        self.rt = D.runtime()
        self.P = D.Program('jac3d')
        self.setup = self.P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])
        self.iterate = self.P.fn('iterate', [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                                 ctypes.c_double)
        self.fetch = self.P.fn('fetch', [ctypes.POINTER(ctypes.c_double)] * 2)
        self.teardown = self.P.fn('teardown', [])
        for name, at in [('dvmh_amd_set_sync_mode_C', [ctypes.c_long]), ('dvmh_amd_set_graph_C', [ctypes.c_long]),
                         ('dvmh_amd_set_kernel_timing_C', [ctypes.c_long]), ('dvmh_amd_reset_kernel_stats_C', []),
                         ('dvmh_amd_kernel_stats_C', [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
                         ('dvmh_amd_tuner_site_pending_C', [ctypes.c_char_p, ctypes.c_long]), ('dvmh_barrier', [])]:
            getattr(self.rt, name).argtypes = at
        self.rt.dvmh_amd_tuner_site_pending_C.restype = ctypes.c_long
        self.rt.dvmh_amd_kernel_stats_C.restype = ctypes.c_long

    def configure(self, form, tuner_mode):
        self.rt.dvmh_amd_tuner_set_mode_C(tuner_mode)
        if form == 'plain':
            self.rt.dvmh_amd_set_sync_mode_C(0)
            self.rt.dvmh_amd_set_graph_C(0)
            self.P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)
        else:
            # One ~300 us kernel per time step: graph replay would save little and
            # would hide the per-kernel events the roofline is measured with, so
            # this workload's chosen form runs its regions eagerly.
            self.rt.dvmh_amd_set_sync_mode_C(1)
            self.rt.dvmh_amd_set_graph_C(0)
            self.P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)

    def pending(self):
        return sum(self.rt.dvmh_amd_tuner_site_pending_C(b'jac3d.cdv', line) for line in STENCIL_LINES)

    def kernel_stats(self):
        launches, ns = 0, 0
        out = (ctypes.c_long * 2)()
        for line in STENCIL_LINES:
            if self.rt.dvmh_amd_kernel_stats_C(b'jac3d.cdv', line, out):
                launches += out[0]
                ns += out[1]
        return launches, ns

    def run_form(self, form, tuner_mode, steps, warmup):
        self.configure(form, tuner_mode)
        self.setup(N, SEED)
        # warm-up: at least `warmup` steps, and until the tuner has measured every candidate
        done = 0
        while done < warmup or (tuner_mode in (1, 2) and self.pending() > 0):
            self.iterate(ITERS, None, None)
            done += 1
            if done > warmup + 40:
                raise RuntimeError('tuner did not settle')
        # one more step so that repeating regions are captured before timing
        if form == 'chosen':
            self.iterate(ITERS, None, None)
        self.rt.dvmh_amd_reset_kernel_stats_C()
        self.rt.dvmh_amd_set_kernel_timing_C(1)
        eps = np.zeros(ITERS)
        s = np.zeros(ITERS)
        self.rt.dvmh_barrier()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.iterate(ITERS, eps.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self.rt.dvmh_barrier()
        t1 = time.perf_counter()
        self.rt.dvmh_amd_set_kernel_timing_C(0)
        launches, ns = self.kernel_stats()
        self.teardown()
        sec = t1 - t0
        value = PTS * ITERS * steps / sec
        kernel_s = ns * 1e-9 / launches if launches else None
        return {'value': value, 'ms_per_step': sec / steps * 1e3, 'kernel_launches': launches,
                'kernel_us': kernel_s * 1e6 if kernel_s else None,
                'achieved_GBps': PTS * B_ALG / kernel_s / 1e9 if kernel_s else None,
                'warmup_steps': done, 'last_eps': float(eps[-1]), 'last_sum': float(s[-1])}


def parity_check(bench, iters=6):
    """Chosen form (graph capture + replay included) against the oracle at
    full size for a few iterations: grid bit-exact, MAX exact, SUM within 1e-12."""
    import helpers as H
    bench.configure('chosen', bench.mode_chosen)
    bench.setup(N, SEED)
    eps = np.zeros(iters)
    s = np.zeros(iters)
    bench.iterate(iters, eps.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    A = np.zeros(N ** 3)
    B = np.zeros(N ** 3)
    bench.fetch(A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), B.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    bench.teardown()
    ref = H.oracle_jac3d(N, iters)
    return {'iterations': iters,
            'grid_bit_exact': bool(H.bits_equal(A, ref['A']) and H.bits_equal(B, ref['B'])),
            'max_exact': bool(np.array_equal(eps, ref['eps'])),
            'sum_max_rel_err': float(np.max(np.abs(s - ref['sum']) / np.abs(ref['sum'])))}


def cpu_baseline(target_seconds=12.0):
    """The oracle's host-handler restatement on all host cores, bounded sample."""
    import helpers as H
    o = H.oracle()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)
    A = np.zeros(N ** 3)
    o.orc_init_grid3d(H.vp(A), N, SEED, 0.0, 1)
    B = A.copy()
    eps = np.zeros(64)
    s = np.zeros(64)
    t0 = time.perf_counter()
    o.orc_jac3d(H.vp(A), H.vp(B), N, 1, H.vp(eps), H.vp(s), cores)
    one = time.perf_counter() - t0
    its = int(max(1, min(4000, target_seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.orc_jac3d(H.vp(A), H.vp(B), N, its, H.vp(eps), H.vp(s), cores)
    sec = time.perf_counter() - t0
    return {'value': PTS * its / sec, 'unit': 'point-updates/s', 'cores': cores, 'kind': 'port',
            'sample': f'{its} iterations of the 384^3 sweep (oracle C/OpenMP host-handler restatement, '
                      f'{sec:.1f} s on {cores} threads)'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--skip-cpu', action='store_true', help='omit the CPU baseline leg')
    ap.add_argument('--skip-parity', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: this build runs on one GPU only (the path is not sharded); --gpus {args.gpus} is not supported',
              file=sys.stderr)
        sys.exit(2)
    bench = Jac3d()
    st = D.stats()
    if not st['gpu_present']:
        raise SystemExit('bench.py: no GPU visible to libdvmh_amd')
    has_model = os.path.exists(MODEL_FILE)
    if has_model:
        os.environ['DVMH_AMD_MODEL'] = MODEL_FILE
    bench.mode_chosen = 2 if has_model else 1
    log('device', D.device_name(), 'CUs', st['cu_count'])
    plain = bench.run_form('plain', 0, args.steps, args.warmup)
    log('plain', plain)
    chosen = bench.run_form('chosen', bench.mode_chosen, args.steps, args.warmup)
    log('chosen', chosen)
    parity = None if args.skip_parity else parity_check(bench)
    log('parity', parity)
    cpu = None if args.skip_cpu else cpu_baseline()
    log('cpu', cpu)
    st = D.stats()
    kernel_s = chosen['kernel_us'] * 1e-6
    achieved = PTS * B_ALG / kernel_s
    out = {
        'metric': METRIC,
        'value': chosen['value'],
        'unit': 'point-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': chosen['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded SplitMix64 interior, zero boundary)',
        'config': {'workload': '3-D 7-point Jacobi, FP64 384^3, 50 iterations per step, REDUCTION(MAX)+REDUCTION(SUM)',
                   'form': 'chosen', 'tuner': 'cost-model' if has_model else 'exhaustive',
                   'graph_replays': st['graph_replays']},
        'roofline': {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK, 'traffic': None,
                     'kernel': 'jac3d stencil loop (loop_jac3d_41/54_kernel)', 'kernel_us': chosen['kernel_us'],
                     'bytes_per_point_update': B_ALG},
        'plain': {'value': plain['value'], 'ms_per_step': plain['ms_per_step'], 'kernel_us': plain['kernel_us'],
                  'frac': (PTS * B_ALG / (plain['kernel_us'] * 1e-6)) / HBM_PEAK if plain['kernel_us'] else None},
        'speedup_vs_plain': chosen['value'] / plain['value'],
        'parity': parity,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
