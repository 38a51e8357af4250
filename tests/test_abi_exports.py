"""The drop-in boundary: libdvmh_amd.so loads without a GPU and exports every
function that include/*.h declares (no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

import dvmh_amd as D

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(REPO, 'include')
HEADERS = ['dvmhlib2.h', 'dvmh_runtime_api.h', 'dvmhlib_stdio.h', 'dvmh_amd.h']


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r'/\*.*?\*/', ' ', text, flags=re.S)
    text = re.sub(r'//[^\n]*', ' ', text)
    text = '\n'.join(l for l in text.split('\n') if not l.lstrip().startswith('#'))
    text = re.sub(r'extern\s+"C"\s*\{', ' ', text)
    # struct/union bodies are not declarations of functions
    text = re.sub(r'\b(struct|union|enum)\s+\w*\s*\{[^}]*\}', ' ', text)
    text = text.replace('}', ' ')
    names = []
    for stmt in text.split(';'):
        stmt = ' '.join(stmt.split())
        if not stmt or stmt.startswith('typedef'):
            continue
        m = re.match(r'^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(', stmt)
        if m:
            names.append(m.group(1))
    return names


def exported_symbols(lib):
    out = subprocess.run(['nm', '-D', '--defined-only', lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize('header', HEADERS)
def test_every_declared_function_is_exported(header):
    names = declared_functions(os.path.join(INCLUDE, header))
    assert len(names) > 0
    syms = exported_symbols(D.lib_path('libdvmh_amd.so'))
    missing = [n for n in names if n not in syms]
    assert not missing, f'{header}: declared but not exported: {missing}'


def test_library_loads_and_reports_no_gpu_here_or_a_gpu_there():
    st = D.stats()
    assert st['gpu_present'] in (0, 1)
    if st['gpu_present']:
        assert 'gfx950' in D.device_name()


def test_entry_points_count():
    """The reference ABI subset (dvmhlib2.h) is the bulk of the boundary."""
    assert len(declared_functions(os.path.join(INCLUDE, 'dvmhlib2.h'))) >= 90


@pytest.mark.parametrize('block,regs,fits', [
    ((1024, 1, 1), 128, 1), ((1024, 1, 1), 129, 0),   # 16 waves: 4 per SIMD, 512 / 4 = 128 registers each
    ((32, 4, 8), 128, 1), ((32, 4, 8), 136, 0),        # the same 1024 threads as a 3-D block
    ((256, 1, 1), 512, 1), ((256, 1, 1), 513, 0),      # one wave per SIMD may take the whole file
    ((64, 1, 1), 512, 1),
    ((512, 1, 1), 256, 1), ((512, 1, 1), 257, 0),      # 257 rounds up to 264 in granules of 8
    ((320, 1, 1), 256, 1), ((320, 1, 1), 257, 0),      # 5 waves: two share a SIMD
    ((1024, 1, 1), 0, 1)])                             # no count given: no constraint
def test_block_register_fit_rule(block, regs, fits):
    """The rule dvmh_loop_cuda_get_config_C applies to the registers per
    thread a handler reports for its kernel (the workgroup's waves over the
    compute unit's four SIMDs, 512 registers per lane each)."""
    r = D.runtime()
    r.dvmh_amd_block_fits_registers_C.restype = ctypes.c_long
    r.dvmh_amd_block_fits_registers_C.argtypes = [ctypes.c_void_p, ctypes.c_long]
    b = (ctypes.c_long * 3)(*block)
    assert r.dvmh_amd_block_fits_registers_C(b, regs) == fits
