"""micrographia_amd: an MI355X-native Rahman molecular-dynamics engine.

The hot path (cell binning, Verlet lists, Lennard-Jones pair kicks and the integer leapfrog) is
hand-written HIP for gfx950 in ``csrc/``, exposed through the C ABI of ``include/rahman_hip.h``
(``librahman_hip.so``).  This package is the Python host side of that ABI, mirroring the
reference's md-core / md engine API.
"""
from ._lib import RahmanError, load
from .core import (
    ONE,
    UNIT,
    Bath,
    BondType,
    ForceMode,
    Membrane,
    Pair,
    Piston,
    State,
    State3,
    System,
    System3,
    Thermostat,
    Wall,
    box_for,
    cube_for,
    fcc,
    fcc_cells,
    fluid,
    force_mode_auto,
    lj_liquid,
    maxwell,
    maxwell2,
    set_temperature,
    thermal_wall,
    to_f64,
    to_fixed,
    verlet3,
    verlet3_with_skin,
    verlet_for,
    verlet_with_skin,
    VERLET_MIN_N,
)
from .engine import CompartmentLedger, GpuEngine, GpuEngine3, Ledger, ListStats, bytes_needed

__all__ = [
    "ONE", "UNIT", "Bath", "BondType", "ForceMode", "Membrane", "Pair", "Piston", "State", "State3",
    "System", "System3", "Thermostat", "Wall", "CompartmentLedger", "Ledger",
    "box_for", "cube_for", "fcc", "fcc_cells", "fluid", "force_mode_auto", "lj_liquid", "maxwell", "maxwell2",
    "set_temperature", "thermal_wall", "to_f64", "to_fixed", "verlet3", "verlet3_with_skin",
    "verlet_for", "verlet_with_skin", "VERLET_MIN_N", "GpuEngine", "GpuEngine3", "ListStats", "RahmanError", "bytes_needed", "load",
]
