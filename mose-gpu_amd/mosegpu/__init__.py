"""Host mirror of the MOSE-GPU residual-path interface on the MI355X C ABI (include/mose_gpu.h)."""
from .core import (BC_NONE, BC_CONNECTION, BC_SYMMETRY, BC_EXTRAPOLATION, BC_FARFIELD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, GC,
                   FARFIELD_DTYPE, INT_CONS, INT_PRIM, LIB_PATH, NTAB, PATH_FAST, PATH_FUSED, PATH_PLAIN, RECORD_DTYPE, Context, DeviceCase, Geometry, Layout, MoseError, Scheme,
                   ThermoTables, device_count, lib, metrics_build, records_tangent, state_hash)
from .core import DeviceChem
from . import cases, chem

__all__ = [
    "BC_NONE", "BC_CONNECTION", "BC_SYMMETRY", "BC_EXTRAPOLATION", "BC_FARFIELD", "FARFIELD_DTYPE", "BC_WALL_ADIABATIC", "BC_WALL_ISOTHERMAL", "GC", "INT_CONS", "INT_PRIM", "LIB_PATH",
    "NTAB", "PATH_FAST", "PATH_FUSED", "PATH_PLAIN", "RECORD_DTYPE", "Context", "DeviceCase", "Geometry", "Layout", "MoseError", "Scheme", "ThermoTables",
    "DeviceChem", "chem", "device_count", "lib", "metrics_build", "records_tangent", "state_hash", "cases",
]
