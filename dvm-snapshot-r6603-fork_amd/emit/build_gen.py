#!/usr/bin/env python3
"""Regenerate gen/<program>_host.c and gen/<program>_hip.hip for every
programs/<program>.loops (the emitter step of the build)."""
import glob
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import hipgen  # noqa: E402

for spec in sorted(glob.glob(os.path.join(PKG, 'programs', '*.loops'))):
    hipgen.emit(spec, os.path.join(PKG, 'gen'))
    print('emitted', os.path.basename(spec))
