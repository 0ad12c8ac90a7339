"""MI355X-native GCMC trial-move energy path (FastMC's insert / delete /
displace / rotate energy terms) as hand-written HIP kernels for gfx950 behind
the C-ABI in include/fastmc.h.

* ``fastmc_amd.system``  -- reader of the reference FIELD / CONFIG / CONTROL files
* ``fastmc_amd.inputs``  -- deterministic synthetic workloads in that format
* ``fastmc_amd.lib``     -- ctypes binding of libfastmc_hip.so
"""
from .system import System, read_system  # noqa: F401

__all__ = ["System", "read_system"]


def handle(*args, **kwargs):
    """Create a device handle (see fastmc_amd.lib.Handle)."""
    from .lib import Handle
    return Handle(*args, **kwargs)
