"""The tuning database across processes (DVMH_AMD_TUNEDB): a process that
tuned its loop sites saves the timings; a fresh process that loads them starts
settled -- the cost-model form picks its configurations from the loaded
timings, times nothing new, and stays bit-exact.  This is the read path the
reference's write-only Scheme.dvmh never had (rts/dvmh/dvmh_rts.cpp:1219-1276,
rts/dvmh/region.cpp:406)."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, 'dvm-snapshot-r6603-fork_amd')

CHILD = r'''
import json, sys
sys.path[:0] = [{pkg!r}, {tests!r}]
import numpy as np
import dvmh_amd as D
import helpers as H
from dvmh_amd import costmodel as CM
mode, iters, save, rounds = int(sys.argv[1]), int(sys.argv[2]), sys.argv[3], int(sys.argv[4])
r = CM.rt()
r.dvmh_amd_set_sync_mode_C.argtypes = [__import__('ctypes').c_long]
assert r.dvmh_amd_model_load_C(CM.MODEL.encode()) >= 0
r.dvmh_amd_tuner_set_mode_C(mode)
r.dvmh_amd_set_sync_mode_C(1)
pending0 = r.dvmh_amd_tuner_pending_C()
# every program run also executes its set-up loop once: runs repeat until no
# site (set-up loops included) has a configuration left to time
for _ in range(rounds):
    j = H.run_jac3d(34, iters)
    s = H.run_sor3d(20, iters)
    a = H.run_adi3d(24, iters)
    pending = r.dvmh_amd_tuner_pending_C()
    if pending == 0:
        break
records = r.dvmh_amd_tuner_save_C(save.encode())
rj, rs, ra = H.oracle_jac3d(34, iters), H.oracle_sor3d(20, iters), H.oracle_adi3d(24, iters)
exact = (H.bits_equal(j['A'], rj['A']) and np.array_equal(j['eps'], rj['eps']) and H.sum_close(j['sum'], rj['sum'])
         and H.bits_equal(s['A'], rs['A']) and np.array_equal(s['eps'], rs['eps'])
         and H.bits_equal(a['A'], ra['A']) and np.array_equal(a['max'], ra['max']) and np.array_equal(a['loc'], ra['loc']))
print('RESULT ' + json.dumps({{'pending_before': int(pending0), 'pending': int(pending), 'records': int(records),
                               'exact': bool(exact), 'gpu_loops': D.stats()['loops_gpu']}}), flush=True)
'''


def _child(mode, iters, save, db=None, rounds=1):
    env = dict(os.environ)
    env.pop('DVMH_AMD_TUNEDB', None)
    if db:
        env['DVMH_AMD_TUNEDB'] = db
    code = CHILD.format(pkg=PKG, tests=os.path.join(REPO, 'tests'))
    print('tunedb child: mode', mode, 'iters', iters, 'db', db, flush=True)
    p = subprocess.run([sys.executable, '-c', code, str(mode), str(iters), save, str(rounds)], env=env, capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    line = [ln for ln in p.stdout.splitlines() if ln.startswith('RESULT ')][-1]
    return json.loads(line[len('RESULT '):])


@pytest.mark.gpu
def test_fresh_process_starts_settled(tmp_path):
    db = str(tmp_path / 'tune.tsv')
    again = str(tmp_path / 'tune_again.tsv')
    # first process: the cost-model form explores its trial sets, then saves
    first = _child(2, 12, db, rounds=12)
    assert first['exact'] and first['pending'] == 0 and first['records'] > 0 and first['gpu_loops'] > 0
    # fresh process, same programs, timings loaded at start-up: nothing explored
    second = _child(2, 4, again, db=db)
    assert second['exact'] and second['gpu_loops'] > 0
    assert second['pending'] == 0
    assert second['records'] == first['records'], 'the fresh process timed configurations again'
    # fixed mode (best recorded configuration only) from the same database
    third = _child(3, 4, str(tmp_path / 'unused.tsv'), db=db)
    assert third['exact'] and third['pending'] == 0
