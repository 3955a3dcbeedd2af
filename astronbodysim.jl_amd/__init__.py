"""MI355X-native direct-summation gravity engine: the (DirectSum, GPU) path of
AstroNbodySim.jl on hand-written HIP kernels for gfx950, behind a C-ABI
(include/nbody_mi355x.h, libnbody_mi355x.so) and this host-side mirror of the
reference's Simulation / run API."""
from . import _native
from ._native import NBError, device_info, version

gpuinfo = device_info   # the reference's gpuinfo(): name, architecture and CU count
from .simulation import (
    AllPairs, BLACKHOLE, BULGE, CPU, DISK, Constants, DirectSum, Euler, GAS, GPU, HALO, Leapfrog,
    MOND1983Milgrom, Newton, QUMOND, STAR,
    Simulation, StarData, Tiled, advance_and_find_timestep, compute_force, compute_potential,
    find_next_sync_point_and_drift, get_all_data, get_local_data, init_timesteps,
    leapfrog_half_kick, numlocal, pack_pos, preprocessdata, run, set_softlen, set_softlen_, softlen, step, total_angular_momentum,
    total_energy, total_kinetic, total_momentum, total_potential,
)
from . import ics
from .restart import loadrestart, saverestart
from .simulation import output, restart
from .snapshot import HeaderGadget2, read_gadget2, write_gadget2
from .ensemble import Ensemble
from . import adaptive  # noqa: F401  (registers the nb_adaptive_* entry points)

__all__ = [
    "NBError", "device_info", "gpuinfo", "version", "saverestart", "loadrestart", "restart", "output",
    "HeaderGadget2", "read_gadget2", "write_gadget2", "Constants", "Newton", "MOND1983Milgrom", "QUMOND",
    "AllPairs", "Tiled", "CPU", "GPU", "DirectSum",
    "Leapfrog", "Euler", "GAS", "HALO", "DISK", "BULGE", "STAR", "BLACKHOLE", "Simulation",
    "StarData", "run", "step", "compute_force", "compute_potential", "total_kinetic",
    "total_potential", "total_energy", "total_momentum", "total_angular_momentum",
    "get_all_data", "get_local_data", "numlocal", "pack_pos", "preprocessdata", "softlen", "set_softlen", "set_softlen_",
    "init_timesteps", "leapfrog_half_kick", "find_next_sync_point_and_drift",
    "advance_and_find_timestep", "ics", "Ensemble",
]
