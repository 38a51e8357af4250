"""Python harness for libdvmh_amd and the converted workload programs.

Loads the in-tree shared libraries through ctypes (plain C ABI, no torch types)
and exposes each program's setup / iterate / fetch / teardown entry points.
Importing never falls back to anything: a missing library raises.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, 'lib')
REPO_DIR = os.path.dirname(PKG_DIR)

SEED_BASE = 0xD0E50000   # workload c uses seed SEED_BASE + c

_libs = {}


class LibraryMissing(RuntimeError):
    pass


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def load(name):
    """Load lib/<name> (e.g. 'libdvmh_amd.so', 'libprog_jac2d.so')."""
    if name in _libs:
        return _libs[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise LibraryMissing(f'{path} is not built; run __graft_entry__.build() first')
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def runtime():
    lib = load('libdvmh_amd.so')
    lib.dvmh_amd_get_stats_C.argtypes = [ctypes.POINTER(ctypes.c_long)]
    lib.dvmh_amd_device_name_C.restype = ctypes.c_char_p
    lib.dvmh_amd_tuner_set_mode_C.argtypes = [ctypes.c_long]
    lib.dvmh_amd_tuner_save_C.argtypes = [ctypes.c_char_p]
    lib.dvmh_amd_tuner_save_C.restype = ctypes.c_long
    lib.dvmh_amd_tuner_load_C.argtypes = [ctypes.c_char_p]
    lib.dvmh_amd_tuner_load_C.restype = ctypes.c_long
    lib.dvmh_amd_tuner_query_C.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    lib.dvmh_amd_tuner_query_C.restype = ctypes.c_long
    lib.dvmh_init_lib_C.argtypes = [ctypes.c_long]
    return lib


def stats():
    lib = runtime()
    out = (ctypes.c_long * 8)()
    lib.dvmh_amd_get_stats_C(out)
    keys = ['loops_host', 'loops_gpu', 'gpu_ns', 'gpu_present', 'cu_count', 'graph_replays', 'batch_groups', 'batch_loops']
    return dict(zip(keys, list(out)))


def reset_stats():
    runtime().dvmh_amd_reset_stats_C()


def device_name():
    return runtime().dvmh_amd_device_name_C().decode()


DEVICE_TYPE_HOST = 1
DEVICE_TYPE_CUDA = 2
REGION_AMD_GRAPH = 256

_dbl_p = ctypes.POINTER(ctypes.c_double)
_long_p = ctypes.POINTER(ctypes.c_long)


def _ptr(arr, ctype=ctypes.c_double):
    if arr is None:
        return None
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


class Program:
    """A converted program library lib/libprog_<name>.so."""

    def __init__(self, name):
        runtime()
        self.name = name
        self.lib = load(f'libprog_{name}.so')
        f = getattr(self.lib, f'{name}_set_targets', None)
        if f is not None:
            f.argtypes = [ctypes.c_long, ctypes.c_int]

    def fn(self, suffix, argtypes, restype=None):
        f = getattr(self.lib, f'{self.name}_{suffix}')
        f.argtypes = argtypes
        f.restype = restype
        return f

    def set_targets(self, device_types, flags=0):
        self.lib.__getattr__(f'{self.name}_set_targets')(device_types, flags)
