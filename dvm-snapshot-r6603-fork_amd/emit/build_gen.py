#!/usr/bin/env python3
"""Regenerate gen/<program>_host.c and gen/<program>_hip.hip for every
programs/<program>.loops (the emitter step of the build), and convert the
C-DVMH sources programs/<program>.cdv of the configurations with the front end
(emit/cdvmh.py) into gen/cdv/<program>_cdv.c plus its loop descriptions,
whose handlers are emitted the same way."""
import glob
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import cdvmh  # noqa: E402
import hipgen  # noqa: E402

# programs/*.cdv that are complete C-DVMH programs: the configurations, and
# relax2d, redall, jacg, wide, redarr, pieces, asyncreg, gauss and seqpart for constructs, reduction kinds,
# handler sizes and data movements they do not use
# (batched.cdv sketches the batched workload's structure and is not
# compilable C)
CDV_PROGRAMS = ['jac2d', 'jac3d', 'sor3d', 'adi3d', 'relax2d', 'redall', 'jacg', 'wide', 'redarr', 'pieces', 'asyncreg', 'gauss', 'seqpart', 'pairs']

for spec in sorted(glob.glob(os.path.join(PKG, 'programs', '*.loops'))):
    hipgen.emit(spec, os.path.join(PKG, 'gen'))
    print('emitted', os.path.basename(spec))

cdv_dir = os.path.join(PKG, 'gen', 'cdv')
for stem in CDV_PROGRAMS:
    cdvmh.convert_file(os.path.join(PKG, 'programs', stem + '.cdv'), cdv_dir, prog=stem + '_cdv')
    hipgen.emit(os.path.join(cdv_dir, stem + '_cdv.loops'), cdv_dir)
    print('converted', stem + '.cdv')
