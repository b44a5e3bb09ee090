"""Python binding of the MI355X BSSN right-hand side (ctypes over the C-ABI
in include/dgr_bssn.h). torch is used only to own device memory; every
computation runs in the HIP library lib/libdgr_bssn.so, which must be built
(dendro-gr_amd/build.py) — there is no fallback.
"""
from .api import (  # noqa: F401
    BssnParams, Plan, NVARS, VAR_NAMES, default_params, lib_path, load, var_pointers,
)
from . import checkpoint, parfile, workloads  # noqa: F401
