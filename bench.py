#!/usr/bin/env python3
"""Benchmark of libdvmh_amd on one MI355X: converted C-DVMH programs running
their PARALLEL / ACROSS loops through the DVMH C interface.

Headline workload: 3-D 7-point Jacobi, FP64 384^3, 50 iterations per step
(each iteration one PARALLEL loop with REDUCTION(MAX) and REDUCTION(SUM)),
inputs resident in HBM.  Three forms run side by side:
  plain       default launch configuration, synchronised after every loop
              (what the reference runtime does);
  exhaustive  the tuner's exhaustive sweep (every candidate configuration
              timed during warm-up, then the fastest), results synchronised
              only when needed;
  chosen      the learned cost model ranks the candidates and only its best
              four plus the default are timed during warm-up.
`value` is the chosen form's point-updates/s.  The roofline uses the stencil
kernel's average duration measured with HIP events on the runtime's stream
during the timed region.  The batched-grid workload (512 uneven 2-D grids) and
the other configurations are reported in the same JSON line, each with a
parity check of its timed form at full size.  The CPU baseline is the DVMH
host path (the product's emitted OpenMP host handlers) timed on a bounded
sample of the headline workload on this machine's host cores.
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, 'dvm-snapshot-r6603-fork_amd')
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(REPO, 'tests'))

import numpy as np  # noqa: E402

import dvmh_amd as D  # noqa: E402
from dvmh_amd import costmodel as CM  # noqa: E402

METRIC = 'FP64 point-updates/s + % HBM roofline; cost-model/batched vs default, bit-exact'
N, ITERS = 384, 50
SEED = D.SEED_BASE + 2
HBM_PEAK = 8.0e12
B_ALG = 16          # bytes per point-update of the stencil loop: read A, write B (8 B each)
PTS = (N - 2) ** 3  # point-updates per iteration
STENCIL_LINES = (41, 54)
TIMING_PERIOD = 5   # headline timed region: HIP events on one stencil launch in 5 per loop site
PMC_SUMMARY = os.path.join(REPO, 'profiles', 'jac3d_384_stencil_pmc.txt')

_dp = ctypes.POINTER(ctypes.c_double)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bind():
    rt = CM.rt()
    for name, at in [('dvmh_amd_set_sync_mode_C', [ctypes.c_long]), ('dvmh_amd_set_graph_C', [ctypes.c_long]),
                     ('dvmh_amd_set_kernel_timing_C', [ctypes.c_long]), ('dvmh_amd_reset_kernel_stats_C', []),
                     ('dvmh_amd_kernel_stats_C', [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
                     ('dvmh_barrier', [])]:
        getattr(rt, name).argtypes = at
    rt.dvmh_amd_kernel_stats_C.restype = ctypes.c_long
    return rt


def kernel_stats(rt, cdv, lines):
    launches, ns = 0, 0
    out = (ctypes.c_long * 2)()
    for line in lines:
        if rt.dvmh_amd_kernel_stats_C(cdv, line, out):
            launches += out[0]
            ns += out[1]
    return launches, ns


def pair_stats(rt):
    out = (ctypes.c_long * 2)()
    rt.dvmh_amd_pair_stats_C(out)
    return list(out)


def set_form(rt, mode, sync, graph=0):
    rt.dvmh_amd_tuner_set_mode_C(mode)
    rt.dvmh_amd_tuner_set_repeats_C(2)
    rt.dvmh_amd_set_sync_mode_C(sync)
    rt.dvmh_amd_set_graph_C(graph)


class Jac3d:
    def __init__(self, rt):
        self.rt = rt
        self.P = D.Program('jac3d')
        self.setup = self.P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])
        self.iterate = self.P.fn('iterate', [ctypes.c_int, _dp, _dp], ctypes.c_double)
        self.fetch = self.P.fn('fetch', [_dp, _dp])
        self.teardown = self.P.fn('teardown', [])
        self.P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)

    def pending(self):
        return sum(self.rt.dvmh_amd_tuner_site_pending_C(b'jac3d.cdv', line) for line in STENCIL_LINES)

    def run_form(self, mode, sync, steps, warmup):
        set_form(self.rt, mode, sync)
        self.setup(N, SEED)
        # warm-up: at least `warmup` steps, and until the tuner has timed its candidates
        done = 0
        while done < warmup or (mode in (1, 2) and self.pending() > 0):
            self.iterate(ITERS, None, None)
            done += 1
            if done > warmup + 120:
                raise RuntimeError('tuner did not settle')
        self.rt.dvmh_amd_reset_kernel_stats_C()
        # the kernel's duration is measured live over the timed region, on one
        # launch in TIMING_PERIOD per loop site (each event pair costs stream time)
        self.rt.dvmh_amd_set_kernel_timing_C(TIMING_PERIOD)
        eps = np.zeros(ITERS)
        s = np.zeros(ITERS)
        self.rt.dvmh_barrier()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.iterate(ITERS, eps.ctypes.data_as(_dp), s.ctypes.data_as(_dp))
        self.rt.dvmh_barrier()
        t1 = time.perf_counter()
        self.rt.dvmh_amd_set_kernel_timing_C(0)
        launches, ns = kernel_stats(self.rt, b'jac3d.cdv', STENCIL_LINES)
        self.teardown()
        sec = t1 - t0
        kernel_s = ns * 1e-9 / launches if launches else None
        return {'value': PTS * ITERS * steps / sec, 'ms_per_step': sec / steps * 1e3, 'kernel_launches': launches,
                'kernel_us': kernel_s * 1e6 if kernel_s else None,
                'achieved_GBps': PTS * B_ALG / kernel_s / 1e9 if kernel_s else None,
                'frac': PTS * B_ALG / kernel_s / HBM_PEAK if kernel_s else None, 'warmup_steps': done}

    def parity(self, mode, iters=6):
        """The given form against the oracle at full size for a few iterations:
        grid bit-exact, MAX exact, SUM within 1e-12 relative."""
        import helpers as H
        set_form(self.rt, mode, 1)
        self.setup(N, SEED)
        eps = np.zeros(iters)
        s = np.zeros(iters)
        self.iterate(iters, eps.ctypes.data_as(_dp), s.ctypes.data_as(_dp))
        A = np.zeros(N ** 3)
        B = np.zeros(N ** 3)
        self.fetch(A.ctypes.data_as(_dp), B.ctypes.data_as(_dp))
        self.teardown()
        ref = H.oracle_jac3d(N, iters)
        return {'iterations': iters,
                'grid_bit_exact': bool(H.bits_equal(A, ref['A']) and H.bits_equal(B, ref['B'])),
                'max_exact': bool(np.array_equal(eps, ref['eps'])),
                'sum_max_rel_err': float(np.max(np.abs(s - ref['sum']) / np.abs(ref['sum'])))}


def batched_leg(rt, iters):
    """512 uneven 2-D Jacobi grids: per-grid launch sequences against one fused
    launch per loop site per iteration; both bit-exact (checked on the
    smallest and largest grid against the oracle)."""
    import helpers as H
    P = D.Program('batched')
    P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)
    count = 512
    setup = P.fn('setup', [ctypes.c_int, ctypes.c_void_p, ctypes.c_ulonglong])
    it = P.fn('iterate', [ctypes.c_int, ctypes.c_int, _dp], ctypes.c_double)
    side = P.fn('side', [ctypes.c_int], ctypes.c_long)
    fetch = P.fn('fetch', [ctypes.c_int, _dp, _dp])
    res = {}
    set_form(rt, 0, 0)
    setup(count, None, D.SEED_BASE + 5)
    sides = np.array([side(g) for g in range(count)])
    pts = int(np.sum((sides - 2) ** 2))
    for name, batched, sync, graph in (('per_grid', 0, 0, 0), ('batched', 1, 1, 1)):
        set_form(rt, 0, sync, graph)
        # (batched: the first iteration runs eagerly, the second is captured as
        # one graph, every later one replays it)
        it(2, batched, None)
        # wall time without events (the per-grid form would pay for thousands of
        # event pairs per iteration), then the kernel time on a second run,
        # issued eagerly (graph replays carry no timing marks)
        replays0 = D.stats()['graph_replays']
        sec = it(iters, batched, None)
        replays = D.stats()['graph_replays'] - replays0
        rt.dvmh_amd_set_graph_C(0)
        rt.dvmh_amd_reset_kernel_stats_C()
        rt.dvmh_amd_set_kernel_timing_C(1)
        it(iters, batched, None)
        rt.dvmh_amd_set_kernel_timing_C(0)
        launches, ns = kernel_stats(rt, b'batched.cdv', (31, 36))
        res[name] = {'value': pts * iters / sec, 'ms_per_iteration': sec / iters * 1e3, 'graph_replays': replays,
                     'kernel_launches': launches,
                     'kernel_ms_per_iteration': ns * 1e-6 / iters,
                     'achieved_GBps': pts * 40 * iters / (ns * 1e-9) / 1e9 if ns else None}
        if batched:
            # each grid's two loops run as one fused pair: at least 24 B per
            # point (A read, B and A written)
            res[name]['achieved_GBps_fused_traffic'] = pts * 24 * iters / (ns * 1e-9) / 1e9 if ns else None
            res[name]['frac_fused_traffic'] = pts * 24 * iters / (ns * 1e-9) / HBM_PEAK if ns else None
    P.fn('teardown', [])()
    # parity: a fresh batched run of 5 iterations in the timed form (eager,
    # captured, then 3 replays), smallest and largest grid
    set_form(rt, 0, 1, 1)
    setup(count, None, D.SEED_BASE + 5)
    eps = np.zeros(5 * count)
    replays0 = D.stats()['graph_replays']
    it(5, 1, eps.ctypes.data_as(_dp))
    parity_replays = D.stats()['graph_replays'] - replays0
    eps = eps.reshape(5, count)
    picks = [int(np.argmin(sides)), int(np.argmax(sides))]
    ok = True
    for r in H.oracle_batched(sides, 5, grids=picks):
        n = int(sides[r['g']])
        A = np.zeros(n * n)
        Bb = np.zeros(n * n)
        fetch(r['g'], A.ctypes.data_as(_dp), Bb.ctypes.data_as(_dp))
        ok = ok and H.bits_equal(A, r['A']) and H.bits_equal(Bb, r['B']) and np.array_equal(eps[:, r['g']], r['eps'])
    P.fn('teardown', [])()
    set_form(rt, 0, 0, 0)
    return {'workload': '512 uneven 2-D Jacobi grids (sides 64..1024 by seed), FP64, 5-point + REDUCTION(MAX) per grid',
            'form': 'batched: each grid\'s stencil and copy + MAX loops fused into one pass, all grids\' pairs in one '
                    'persistent kernel (plus one over the tile-edge points) per iteration, the time step replayed as one graph',
            'parity_graph_replays': parity_replays,
            'points_per_iteration': pts, 'iterations': iters, 'bytes_per_point_update': 40,
            'per_grid': res['per_grid'], 'batched': res['batched'],
            'speedup': res['batched']['value'] / res['per_grid']['value'], 'parity_bit_exact': bool(ok)}


def other_parity(rt, name, n, iters, chain=False):
    """The chosen form (cost-model configuration settled by the timed legs,
    results synchronised only when needed; repeating regions replayed as HIP
    graphs, or SOR's sweeps run as one chain) at the size and iteration count
    it is timed, from a fresh setup, against the oracle: grids bit-exact, MAX /
    MAXLOC values and indices exact."""
    import helpers as H
    set_form(rt, 2, 1, 0 if chain else 1)
    flags = 0 if chain else D.REGION_AMD_GRAPH
    st0 = D.stats()
    if name == 'jac2d':
        res, ref = H.run_jac2d(n, iters, flags=flags), H.oracle_jac2d(n, iters)
        ok = {'grid_bit_exact': bool(H.bits_equal(res['A'], ref['A']) and H.bits_equal(res['B'], ref['B'])),
              'max_exact': bool(np.array_equal(res['eps'], ref['eps']))}
    elif name == 'sor3d':
        res = H.run_sor3d_chain(n, iters, iters) if chain else H.run_sor3d(n, iters, flags=flags)
        ref = H.oracle_sor3d(n, iters)
        ok = {'grid_bit_exact': bool(H.bits_equal(res['A'], ref['A'])),
              'max_exact': bool(np.array_equal(res['eps'], ref['eps']))}
    else:
        res, ref = H.run_adi3d(n, iters, flags=flags), H.oracle_adi3d(n, iters)
        ok = {'grid_bit_exact': bool(H.bits_equal(res['A'], ref['A'])),
              'max_exact': bool(np.array_equal(res['max'], ref['max'])),
              'loc_exact': bool(np.array_equal(res['loc'], ref['loc']))}
    st1 = D.stats()
    ok['iterations'] = iters
    replays = st1['graph_replays'] - st0['graph_replays']
    if chain:
        ok['form'] = 'chain of %d sweeps' % iters
    elif replays:
        ok['form'] = 'deferred region replayed as a graph'
    else:
        # a region whose graph is below the replay threshold runs eagerly
        ok['form'] = 'deferred region launched eagerly (its graph is too small to pay for a graph launch)'
    ok['graph_replays'] = replays
    ok['batch_groups'] = st1['batch_groups'] - st0['batch_groups']
    set_form(rt, 0, 0, 0)
    return ok


def chosen_layouts(rt, w, chain=False):
    """Kernel variant and device layout the tuner kept at each loop site of the
    workload (from the tuner's query; transform 0 = natural layout)."""
    out = {}
    q = (ctypes.c_long * 13)()
    f = rt.dvmh_amd_tuner_query_chain_C if chain else rt.dvmh_amd_tuner_query_C
    f.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    f.restype = ctypes.c_long
    ts = rt.dvmh_amd_tuner_transform_stats_C
    ts.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    ts.restype = ctypes.c_long
    t = (ctypes.c_long * 4)()
    for line in w.lines:
        if f(w.cdv, line, q) > 0:
            out[str(line)] = {'block': [q[2], q[3], q[4]], 'variant': q[5 + 2], 'tile_a': q[5 + 4], 'tile_b': q[5 + 5],
                              'chunk': q[5 + 3], 'transform': q[5 + 6]}
            if not chain and ts(w.cdv, line, t) > 0 and t[0] > 0:
                # the site's layout-transform candidates: how many the tuner timed
                # (the model's trial set) and the best of them against the best
                # natural-layout one (transform round trip included)
                out[str(line)]['transform_candidates'] = {'count': t[0], 'timed': t[1],
                                                          'best_us': t[2] / 1e3 if t[2] >= 0 else None,
                                                          'best_natural_us': t[3] / 1e3 if t[3] >= 0 else None}
    # a fused pair's own site (first loop's line)
    fp = rt.dvmh_amd_tuner_query_pair_C
    fp.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    fp.restype = ctypes.c_long
    for line in w.lines:
        if not chain and fp(w.cdv, line, q) > 0:
            out[f'{line} (fused pair)'] = {'block': [q[2], q[3], q[4]], 'tile_a': q[5 + 4], 'chunk': q[5 + 3], 'xcd': q[5 + 1],
                                           'us': q[1] / 1e3}
    return out


def layout_transforms(reps=5):
    """The layout-transform kernels (csrc/kernels/transform.hip) at the shapes
    the ACROSS configurations would re-lay out: the ADI 256^3 array (exchange of
    the two fastest axes: tiled transpose; axis rotation: row gather) and the
    SOR 320^3 array (diagonal planes).  Each transform reads and writes the
    array once: algorithmic bytes = 2 x 8 B x elements."""
    rt = D.runtime()
    f = rt.dvmh_amd_layout_transform_seconds_C
    f.argtypes = [ctypes.c_long] * 5 + [ctypes.POINTER(ctypes.c_double)]
    f.restype = ctypes.c_long
    o = (ctypes.c_double * 2)()
    out = {}
    for key, which, n, desc in (('transpose_256', 0, 256, 'exchange of the two fastest axes, FP64 256^3 (transposeTiles)'),
                                ('diagonal_320', 1, 320, 'natural <-> diagonal planes, FP64 320^3 (diagonalTiles)'),
                                ('rotate_256', 2, 256, 'axis rotation, FP64 256^3 (permuteGather)')):
        if f(which, n, n, n, reps, o) != 0:
            continue
        nbytes = 2.0 * 8 * n ** 3
        out[key] = {'what': desc, 'forward_us': o[0] * 1e6, 'backward_us': o[1] * 1e6,
                    'forward_GBps': nbytes / o[0] / 1e9, 'backward_GBps': nbytes / o[1] / 1e9,
                    'frac_forward': nbytes / o[0] / HBM_PEAK, 'frac_backward': nbytes / o[1] / HBM_PEAK}
    return out


def time_form(rt, name, n, iters, mode, sync, graph):
    """One form of a workload: tuner settled first, then the wall time of a run
    without timing events, then the kernels' time from HIP events on a second
    run whose launches are issued eagerly (graph replays carry no timing
    marks), so that `frac` is the kernels' algorithmic bytes over their own
    time."""
    from dvmh_amd import workloads as W
    set_form(rt, mode, sync, graph)
    w = W.ALL[name](n, flags=D.REGION_AMD_GRAPH if graph else 0)
    w.setup()
    for _ in range(200):
        w.run(max(2, iters // 10) if name != 'sor3d_chain' else iters)
        if mode == 0 or sum(rt.dvmh_amd_tuner_site_pending_C(w.cdv, ln) for ln in w.lines) == 0:
            break
    w.run(3 if name != 'sor3d_chain' else iters)   # settled: a repeating region is captured here, before timing
    sec = w.run(iters)
    # the same form without capture (regions still deferred, so loop pairs
    # still run fused), every launch between timing marks
    rt.dvmh_amd_set_graph_C(2 if graph else 0)
    rt.dvmh_amd_reset_kernel_stats_C()
    rt.dvmh_amd_set_kernel_timing_C(1)
    p0 = pair_stats(rt)
    w.run(iters)
    pairs = pair_stats(rt)[1] - p0[1]
    rt.dvmh_amd_set_kernel_timing_C(0)
    rt.dvmh_amd_set_graph_C(graph)
    launches, ns = kernel_stats(rt, w.cdv, w.lines)
    layouts = chosen_layouts(rt, w, chain=name == 'sor3d_chain') if mode == 2 else None
    w.teardown()
    upd, byt = w.points_per_iter(), w.bytes_per_iter()
    r = {'value': upd * iters / sec, 'ms': sec * 1e3, 'ms_per_iteration': sec * 1e3 / iters,
         'kernel_ms': ns * 1e-6 if launches else None, 'kernel_launch_marks': launches,
         'frac': byt * iters / (ns * 1e-9) / HBM_PEAK if ns else None,
         'frac_wall': byt * iters / sec / HBM_PEAK,
         'point_updates_per_iteration': upd, 'bytes_per_iteration': byt}
    if pairs:
        # fused loop pairs: the kernels move at least the fused step's bytes
        # (each array read once and written once per point)
        fb = w.fused_bytes_per_iter()
        r['fused_pairs_per_iteration'] = pairs / iters
        r['fused_bytes_per_iteration'] = fb
        r['frac_fused_traffic'] = fb * iters / (ns * 1e-9) / HBM_PEAK if ns else None
    if layouts is not None:
        r['sites'] = layouts
    return r


def other_configs(rt, parity=True, only=None):
    """The remaining configurations, plain (default configuration, per-loop
    launches) against the chosen form (cost model, results synchronised only
    when needed, repeating time steps replayed as graphs; SOR's sweeps chained
    30 at a time), each chosen form then checked against the oracle for the
    whole timed run."""
    specs = [('jac2d', 2048, 100, 'C1 2-D Jacobi 2048^2 x 100 (two loops per iteration)'),
             ('sor3d', 320, 30, 'C3 SOR 320^3 x 30 (ACROSS on three axes, REDUCTION(MAX))'),
             ('adi3d', 256, 20, 'C4 ADI 256^3 x 20 (six line sweeps + MAXLOC; 6 x 255 x 256^2 point-updates per step)')]
    out = {}
    for name, n, iters, desc in specs:
        if only is not None and name not in only:
            continue
        r = {'workload': desc}
        r['plain'] = time_form(rt, name, n, iters, 0, 0, 0)
        if name == 'sor3d':
            # per-sweep chosen form (each sweep its own launch sequence), then the chain
            r['chosen_per_sweep'] = time_form(rt, name, n, iters, 2, 1, 1)
            r['chosen'] = time_form(rt, 'sor3d_chain', n, iters, 2, 1, 0)
            r['chosen']['form'] = 'cost model per loop, the 30 sweeps as one chain (chain tile shape tuned per sweep)'
        else:
            r['chosen'] = time_form(rt, name, n, iters, 2, 1, 1)
        r['speedup'] = r['chosen']['value'] / r['plain']['value']
        if parity:
            r['parity'] = other_parity(rt, name, n, iters, chain=name == 'sor3d')
            log(name, 'parity', r['parity'])
        out[name] = r
        log(name, {k: (v if not isinstance(v, dict) else {kk: vv for kk, vv in v.items() if kk != 'sites'}) for k, v in r.items()})
    set_form(rt, 0, 0, 0)
    return out


def cpu_baseline(target_seconds=12.0):
    """The DVMH host path on the host cores, on a bounded sample of the
    headline workload: the jac3d program with its regions run on the host,
    i.e. through the emitted host handlers (cdvmh's host-handler shape, OpenMP
    over the first loop axis, the loop body's own arithmetic and per-thread
    reduction partials merged as red_post does).  The oracle's restatement on
    the same cores is reported beside it as a second figure."""
    import helpers as H
    rt = D.runtime()
    rt.dvmh_amd_host_slots_C.restype = ctypes.c_long
    threads = int(rt.dvmh_amd_host_slots_C())
    P = D.Program('jac3d')
    setup = P.fn('setup', [ctypes.c_long, ctypes.c_ulonglong])
    iterate = P.fn('iterate', [ctypes.c_int, _dp, _dp], ctypes.c_double)
    teardown = P.fn('teardown', [])
    P.set_targets(D.DEVICE_TYPE_HOST, 0)
    D.reset_stats()
    setup(N, SEED)
    one = iterate(2, None, None) / 2
    its = int(max(2, min(4000, target_seconds / max(one, 1e-4))))
    sec = iterate(its, None, None)
    teardown()
    st = D.stats()
    P.set_targets(D.DEVICE_TYPE_HOST | D.DEVICE_TYPE_CUDA, 0)
    assert st['loops_gpu'] == 0 and st['loops_host'] > 0, st
    # second figure: the oracle (test infrastructure) on the same threads
    o = H.oracle()
    A = np.zeros(N ** 3)
    o.orc_init_grid3d(H.vp(A), N, SEED, 0.0, 1)
    B = A.copy()
    oits = max(2, its // 2)
    eps = np.zeros(oits)
    sm = np.zeros(oits)
    t0 = time.perf_counter()
    o.orc_jac3d(H.vp(A), H.vp(B), N, oits, H.vp(eps), H.vp(sm), threads)
    osec = time.perf_counter() - t0
    return {'value': PTS * its / sec, 'unit': 'point-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'{its} iterations of the 384^3 Jacobi sweep through the emitted host handlers '
                      f'(jac3d program, regions on the host; {sec:.1f} s on {threads} OpenMP threads)',
            'cpu_model': cpu_model(),
            'oracle_restatement': {'value': PTS * oits / osec, 'iterations': oits, 'seconds': osec,
                                   'note': 'oracle C/OpenMP restatement (compensated SUM), same threads'}}


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def measured_traffic():
    """HBM bytes per stencil launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration it states)."""
    if not os.path.exists(PMC_SUMMARY):
        return None
    m = re.search(r'traffic_bytes_per_launch:\s*([0-9.eE+]+)', open(PMC_SUMMARY).read())
    return float(m.group(1)) if m else None


def _stream_rt():
    rt = D.runtime()
    for name in ('dvmh_amd_stream_copy_seconds_C', 'dvmh_amd_stream_read_seconds_C'):
        f = getattr(rt, name)
        f.argtypes = [ctypes.c_long, ctypes.c_long]
        f.restype = ctypes.c_double
    return rt


def calibrate_copy(nbytes=1 << 30):
    """Known-traffic dispatches for scaling FETCH_SIZE / WRITE_SIZE of a
    profiling run: two launches of the runtime's streaming copy (nbytes read
    and nbytes written each) and two of its read-only pass (nbytes read)."""
    rt = _stream_rt()
    rt.dvmh_amd_stream_copy_seconds_C(nbytes, 1)
    rt.dvmh_amd_stream_read_seconds_C(nbytes, 1)
    log('calibration: 2 streaming copies and 2 read-only passes of', nbytes, 'bytes done')


def copy_rate(nbytes=PTS * 8, reps=10):
    """Rates of the runtime's own streaming kernels on this part (GB/s): a copy
    that reads and writes each byte once (counting both), and a read-only
    pass; the practical ceilings of a read-once/write-once kernel."""
    rt = _stream_rt()
    c = rt.dvmh_amd_stream_copy_seconds_C(nbytes, reps)
    r = rt.dvmh_amd_stream_read_seconds_C(nbytes, reps)
    return 2.0 * nbytes / c / 1e9, nbytes / r / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--skip-cpu', action='store_true', help='omit the CPU baseline leg')
    ap.add_argument('--skip-parity', action='store_true')
    ap.add_argument('--skip-others', action='store_true', help='omit the batched-grid and other-config legs')
    ap.add_argument('--chosen-only', action='store_true',
                    help='run only the chosen form (for profiling runs); prints the JSON line without the other legs')
    ap.add_argument('--others-only', action='store_true',
                    help='run only the batched-grid and other-config legs (for profiling runs); prints their JSON')
    ap.add_argument('--legs', default=None,
                    help='with --others-only: comma-separated legs to run (batched, jac2d, sor3d, adi3d, transforms)')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='also run the streaming copy and read-only kernels on 1 GiB (scale reference for HBM counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: this build runs on one GPU only (the path is not sharded); --gpus {args.gpus} is not supported',
              file=sys.stderr)
        sys.exit(2)
    rt = bind()
    if not os.path.exists(CM.MODEL) or rt.dvmh_amd_model_load_C(CM.MODEL.encode()) < 0:
        raise SystemExit(f'bench.py: cannot load the trained cost model {CM.MODEL}')
    st = D.stats()
    if not st['gpu_present']:
        raise SystemExit('bench.py: no GPU visible to libdvmh_amd')
    log('device', D.device_name(), 'CUs', st['cu_count'])
    if args.calibrate_copy:
        calibrate_copy()
    if args.others_only:
        legs = set(args.legs.split(',')) if args.legs else {'batched', 'jac2d', 'sor3d', 'adi3d', 'transforms'}
        batched = batched_leg(rt, 20) if 'batched' in legs else None
        log('batched', batched)
        others = other_configs(rt, only=legs)
        log('others', others)
        transforms = layout_transforms() if 'transforms' in legs else None
        log('layout transforms', transforms)
        print(json.dumps({'batched_grids': batched, 'other_configs': others, 'layout_transforms': transforms}), flush=True)
        return
    bench = Jac3d(rt)
    plain = exhaustive = None
    if not args.chosen_only:
        plain = bench.run_form(0, 0, args.steps, args.warmup)
        log('plain', plain)
        rt.dvmh_amd_tuner_clear_C()
        exhaustive = bench.run_form(1, 1, args.steps, args.warmup)
        log('exhaustive', exhaustive)
        rt.dvmh_amd_tuner_clear_C()
    chosen = bench.run_form(2, 1, args.steps, args.warmup)
    log('chosen', chosen)
    parity = None if (args.skip_parity or args.chosen_only) else bench.parity(2)
    log('parity', parity)
    batched = others = transforms = None
    if not (args.skip_others or args.chosen_only):
        batched = batched_leg(rt, 20)
        log('batched', batched)
        others = other_configs(rt)
        log('others', others)
        transforms = layout_transforms()
        log('layout transforms', transforms)
    cpu = None if (args.skip_cpu or args.chosen_only) else cpu_baseline()
    log('cpu', cpu)
    kernel_s = chosen['kernel_us'] * 1e-6
    achieved = PTS * B_ALG / kernel_s
    copy_gbps, read_gbps = copy_rate()
    log('device copy', copy_gbps, 'GB/s; read-only pass', read_gbps, 'GB/s')
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
                   'form': 'chosen (learned cost model: top-4 + default timed)'},
        'roofline': {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK, 'traffic': measured_traffic(),
                     'kernel': 'jac3d stencil loop kernels (loop_jac3d_41/54, the configuration the tuner kept)', 'kernel_us': chosen['kernel_us'],
                     'bytes_per_point_update': B_ALG, 'point_updates_per_launch': PTS,
                     'device_copy_GBps': copy_gbps, 'device_read_GBps': read_gbps,
                     'frac_of_device_copy': achieved / 1e9 / copy_gbps,
                     'ceiling': 'streaming copy kernel of the runtime (16-byte lanes, 4 loads in flight), same bytes'},
        'plain': plain,
        'exhaustive': exhaustive,
        'speedup_vs_plain': chosen['value'] / plain['value'] if plain else None,
        'kernel_speedup_vs_plain': plain['kernel_us'] / chosen['kernel_us'] if plain else None,
        'cost_model_vs_exhaustive_kernel': chosen['kernel_us'] / exhaustive['kernel_us'] if exhaustive else None,
        'parity': parity,
        'batched_grids': batched,
        'other_configs': others,
        'layout_transforms': transforms,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
