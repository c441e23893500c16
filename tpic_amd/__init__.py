"""tpic_amd: MI355X-native ECSIM-cycle hot path (mover, current and mass
deposition, per-cell sort, implicit Maxwell GMRES) behind the C-ABI of
include/tpic_mi.h. The native library is tpic_amd/lib/libtpic_mi.so."""
from .runtime import (  # noqa: F401
    FIELD_B, FIELD_BNODE, FIELD_E, FIELD_ETHETA, FIELD_J, FIELD_MASS, FIELD_RHS, MASS_BF16X3, MASS_FP32, MASS_FP64,
    Simulation,
    TpicError,
    library_path, load_library,
)
from . import cases  # noqa: F401

__all__ = ["Simulation", "TpicError", "load_library", "library_path", "cases"]
