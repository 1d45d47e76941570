"""rptunnel -- host side of the MI355X ring-polymer hot path of pimd-tunneling.

Thin Python mirror of the reference's Fortran interfaces (mcmod_mass, verletint,
instantonmod) over the C ABI of libpimdtunnel.so (include/rp_hip.h).  All compute
runs in the HIP library; this package only marshals arrays and drives loops the way
the reference's drivers do.
"""
from . import specs
from ._lib import RpError, init, lib
from .pes import Pes
from .verlet import Chain
from .batch import Batch
from .instanton import Ring

__all__ = ["specs", "RpError", "init", "lib", "Pes", "Chain", "Batch", "Ring"]
