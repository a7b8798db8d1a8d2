"""freeflow_amd: MI355X (gfx950) HOME-LBM fluid step with moving-boundary coupling.

The compute path is libfreeflow_lbm.so (hand-written HIP behind the C ABI in
include/freeflow_lbm.h).  `freeflow_amd.capi` is a ctypes binding of that ABI;
`fsi_simulator` (built into freeflow_amd/lib) is the pybind11 drop-in for
FreeFlow's Python module.  There is no CPU fallback.
"""
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")


def enable_dropin():
    """Make `import fsi_simulator` resolve to this package's drop-in module."""
    if LIB_DIR not in sys.path:
        sys.path.insert(0, LIB_DIR)
    import fsi_simulator  # noqa: F401

    return fsi_simulator
