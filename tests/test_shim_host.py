"""Host-side behaviour of the reference-signature shims that needs no GPU:
the library rejects a block beyond the unzipped arrays before any device
call, and the domain setter keeps what it was given."""
import ctypes as C

import numpy as np

import dgr
from dgr import workloads as W


def test_shim_domain_round_trip_and_bounds_check():
    L = dgr.load()
    d3 = C.c_double * 3
    L.dgr_shim_set_domain.argtypes = [d3, d3, d3, d3, C.c_uint64]
    L.dgr_shim_set_domain(d3(-4, -4, -4), d3(4, 4, 4), d3(0, 0, 0), d3(256, 256, 256), 1000)

    class Dom(C.Structure):
        _fields_ = [("compd_min", C.c_double * 3), ("compd_max", C.c_double * 3),
                    ("octree_min", C.c_double * 3), ("octree_max", C.c_double * 3),
                    ("szpdof", C.c_uint64), ("set", C.c_int32)]

    L.dgr_shim_domain.restype = C.POINTER(Dom)
    d = L.dgr_shim_domain().contents
    assert list(d.compd_min) == [-4, -4, -4] and list(d.octree_max) == [256, 256, 256]
    assert d.szpdof == 1000 and d.set == 1
    # a 19^3 block does not fit in 1000 points per variable
    t = W.single_block_19()
    L.dgr_shim_rhs_block.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
    ptrs = (C.c_void_p * 24)(*([1] * 24))
    rc = L.dgr_shim_rhs_block(np.ascontiguousarray(t).ctypes.data, ptrs, ptrs, 0.0)
    assert rc == -1 and b"szpdof" in L.dgr_last_error()
    built, calls = C.c_int64(), C.c_int64()
    L.dgr_shim_cache_stats(C.byref(built), C.byref(calls))
    assert built.value == 0 and calls.value == 0
    L.dgr_shim_release()
