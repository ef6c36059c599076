"""MI355X-native (gfx950) implementation of the per-step terrain pipeline of
MarWo22/master_thesis: plate advection with a (cell, plate) key sort,
collision resolution, uplift, thermal and hydraulic erosion and
connected-component labelling of plate masks, behind the C ABI of
libtectonic.so (include/tectonic.h)."""
from . import _abi, configs  # noqa: F401
from ._abi import KernelSettings, PlateData, SimConfig, StepStats  # noqa: F401

__all__ = ["_abi", "configs", "KernelSettings", "SimConfig", "PlateData", "StepStats"]
