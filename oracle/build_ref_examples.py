"""Builds the reference's own C-DVMH example programs against this runtime:
each `examples/<stem>.cdv` under /root/reference is converted by the front end
(emit/cdvmh.py), its handlers emitted (emit/hipgen.py), compiled and linked
against libdvmh_amd.  Test infrastructure only: the intermediate files live in
a temporary directory that is removed; only the executables are written, to
oracle/_ref/examples/<name> (kept out of the repository's history), so that
the GPU tests can run the reference's programs where the reference tree is
absent.  Nothing here is part of the runtime.

    python3 oracle/build_ref_examples.py          # every example in EXAMPLES
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

REF = '/root/reference/dvm_sys/examples'
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, 'dvm-snapshot-r6603-fork_amd')
OUT = os.path.join(HERE, '_ref', 'examples')

# name -> (example stem, {macro: value} rewritten before conversion)
EXAMPLES = {
    'jac2d': ('jac2d', None), 'jacross': ('jacross', None), 'jac3d': ('jac3d', None), 'sor': ('sor', None),
    'redb_c': ('redb_c', None), 'testcdvm': ('testcdvm', None), 'gauss_c': ('gauss_c', None),
    'gauss_wb': ('gauss_wb', None), 'mgrid': ('mgrid', None), 'tsk_lp': ('tsk_lp', None),
    'adi20': ('adi', {'nx': 20, 'ny': 20, 'nz': 20}), 'adi': ('adi', None),
    'jac3d_perf': ('jac3d_perf', None), 'jac2d_perf': ('jac2d_perf', None),
}


def build(stem, exe, work, sizes=None, rpath_origin=False):
    """Converts examples/<stem>.cdv (with `#define macro N` lines rewritten per
    `sizes`) in `work` and links the executable `exe`."""
    sys.path.insert(0, os.path.join(PKG, 'emit'))
    import cdvmh
    import hipgen
    src = os.path.join(REF, stem + '.cdv')
    if sizes:
        with open(src) as f:
            text = f.read()
        for name, val in sizes.items():
            text, n = re.subn(r'(#define\s+%s\s+)\d+' % name, r'\g<1>%d' % val, text)
            if n != 1:
                raise ValueError(f'{stem}: macro {name} not found once')
        src = os.path.join(work, stem + '.cdv')
        with open(src, 'w') as f:
            f.write(text)
    cdvmh.convert_file(src, work, prog=stem)
    hipgen.emit(os.path.join(work, stem + '.loops'), work)
    inc = ['-I' + os.path.join(REPO, 'include'), '-I' + os.path.join(PKG, 'programs')]
    cc = ['gcc', '-O2', '-fPIC', '-fopenmp', '-ffp-contract=off', '-D_DVMH=1'] + inc
    objs = []
    for s in (f'{stem}.c', f'{stem}_host.c'):
        o = os.path.join(work, s + '.o')
        subprocess.run(cc + ['-c', os.path.join(work, s), '-o', o], check=True, timeout=120)
        objs.append(o)
    o = os.path.join(work, stem + '_hip.o')
    subprocess.run(['/opt/rocm/bin/hipcc', '--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-ffp-contract=off',
                    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-D__HIP_PLATFORM_AMD__',
                    '-I' + os.path.join(PKG, 'csrc', 'rt'), '-I' + os.path.join(PKG, 'csrc', 'kernels')] + inc +
                   ['-c', os.path.join(work, stem + '_hip.hip'), '-o', o], check=True, timeout=300,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    objs.append(o)
    lib = os.path.join(PKG, 'lib')
    # executables kept under oracle/_ref find the runtime relative to themselves
    rp = '$ORIGIN/' + os.path.relpath(lib, os.path.dirname(exe)) if rpath_origin else lib
    subprocess.run(['g++', '-fopenmp', '-o', exe] + objs + ['-L' + lib, '-ldvmh_amd', '-L/opt/rocm/lib', '-lamdhip64',
                                                           '-Wl,-rpath,' + rp, '-Wl,-rpath,/opt/rocm/lib'],
                   check=True, timeout=120)


def build_all():
    if not os.path.isdir(REF):
        print('build_ref_examples: the reference tree is not present; nothing to build')
        return 0
    os.makedirs(OUT, exist_ok=True)
    for name, (stem, sizes) in EXAMPLES.items():
        work = tempfile.mkdtemp(prefix='refex_')
        try:
            build(stem, os.path.join(OUT, name), work, sizes, rpath_origin=True)
        finally:
            shutil.rmtree(work, ignore_errors=True)
    print(f'build_ref_examples: {len(EXAMPLES)} executables in {OUT}')
    return len(EXAMPLES)


if __name__ == '__main__':
    build_all()
