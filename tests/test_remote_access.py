"""REMOTE_ACCESS outside parallel loops through the C ABI
(dvmh_remote_access_C, the call cdvmh emits for a remote_access block,
cdvmh/src/converter_main.cpp:1250-1300): with one process every element is
local, so the buffer the call describes is a read-only alias of the array's
host copy, made current for the elements named.  Reading through the alias
gives the array's values; deleting the alias leaves the array intact."""
import ctypes

import numpy as np

import dvmh_amd as D

L = ctypes.c_long
RT_DOUBLE = 4   # include/dvmhlib_const.h


def _rt():
    rt = D.runtime()
    rt.dvmh_get_element_addr_C.restype = ctypes.c_void_p
    for name in ('dvmh_distribution_block_C', 'dvmh_distribution_replicated_', 'dvmh_alignment_linear_C'):
        getattr(rt, name).restype = L
    return rt


def _elem(rt, hdr, i, j):
    return ctypes.c_double.from_address(rt.dvmh_get_element_addr_C(hdr, L(2), L(i), L(j)))


def test_remote_access_alias():
    rt = _rt()
    rt.dvmh_init_lib_C(L(0))
    n, m = 5, 7
    hdr = (L * 64)()
    rt.dvmh_array_create_C(hdr, L(2), L(-RT_DOUBLE), L(n), L(0), L(0), L(m), L(0), L(0))
    rt.dvmh_distribute_C(hdr, L(2), L(rt.dvmh_distribution_block_C(L(1))), L(rt.dvmh_distribution_replicated_()))
    vals = np.arange(n * m, dtype=np.float64).reshape(n, m) * 0.5 + 1.0
    for i in range(n):
        for j in range(m):
            _elem(rt, hdr, i, j).value = vals[i, j]
    rma = (L * 64)()
    # row 3, every column: a constant rule, then a replicated one
    rt.dvmh_remote_access_C(rma, hdr, L(2), L(rt.dvmh_alignment_linear_C(L(0), L(0), L(3))),
                            L(rt.dvmh_alignment_linear_C(L(-1), L(0), L(0))))
    assert rma[0] != 0 and rma[0] != hdr[0]
    got = [_elem(rt, rma, 3, j).value for j in range(m)]
    assert got == list(vals[3])
    rt.dvmh_delete_object_(rma)
    assert [_elem(rt, hdr, 3, j).value for j in range(m)] == list(vals[3])
    rt.dvmh_delete_object_(hdr)
