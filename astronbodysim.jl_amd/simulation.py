"""Host-side mirror of AstroNbodySim's Simulation / run API for the
(DirectSum, GPU) path, driving libnbody_mi355x.so through its C-ABI.

Names, keyword arguments, call order and error behaviour follow the
reference (paths relative to the AstroNbodySim.jl repository root):

  Simulation(data; kw...)            src/base/Config.jl:968-1172 (GPU branch :1072-1089)
  run(sim)                           src/run.jl:46-184
  step(sim)                          src/directsumgpu/timestep.jl:216-224
  compute_force(sim[, pos, h])       src/base/Gravity.jl:1-8
  compute_potential(sim[, pos, h])   src/base/Potential.jl:1-24
  total_kinetic / total_potential / total_energy / total_momentum /
  total_angular_momentum             src/base/Energy.jl:13-100
  get_all_data / get_local_data      src/base/Config.jl:1182-1235
  softlen / set_softlen!             src/tools/SofteningLength.jl:10-93
  output(sim)  (Gadget-2 snapshots)  src/base/Output.jl:9-63
  analysis.csv / timing.csv          src/base/Logging.jl:1-21, src/base/Analyse.jl:7-22,
                                     src/base/Timing.jl:32-37

Quantities are unitless (the reference's ``units = nothing`` path); a caller
with Unitful data strips units first, as preprocessdata does
(src/run.jl:9-11).  Only ``device = GPU()`` is served: the CPU path is the
reference's own, and requesting it raises.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _native as nat


# ---------------------------------------------------------------------------
# Traits  (src/base/Traits.jl:1-48; CPU/GPU/Leapfrog/Euler live in AstroSimBase)
# ---------------------------------------------------------------------------
class DeviceType:
    pass


class CPU(DeviceType):
    pass


class GPU(DeviceType):
    pass


class GPUAlgorithm:
    pass


class AllPairs(GPUAlgorithm):
    pass


class Tiled(GPUAlgorithm):
    pass


class Gravity:
    pass


class DirectSum(Gravity):
    pass


class TimeIntegration:
    pass


class Leapfrog(TimeIntegration):
    pass


class Euler(TimeIntegration):
    pass


# Collection codes: Int(c) of the reference's Collection enum, 1..6
GAS, HALO, DISK, BULGE, STAR, BLACKHOLE = 1, 2, 3, 4, 5, 6


# ---------------------------------------------------------------------------
# Configuration  (src/base/Config.jl)
# ---------------------------------------------------------------------------
@dataclass
class ConstantTimestep:
    dt: float


@dataclass
class AdaptiveTimestep:
    MinStep: float
    MaxStep: float
    TimeBase: int
    TimeInterval: float
    ErrTolTimestep: float


@dataclass
class TimeConfig:
    Begin: float
    End: float
    BetweenSnapshots: float
    BetweenSnapshotsInt: int
    step: object
    algorithm: TimeIntegration


class GravityModel:
    """Gravity-model traits (src/base/Config.jl:110)."""


class Newton(GravityModel):
    pass


class MOND1983Milgrom(GravityModel):
    """Newtonian forces rescaled by nu(|a|/ACC0) after every force evaluation
    (src/mond/milgrom1983.jl:33-47, src/base/PostProcessing/Force.jl:41-44)."""


class QUMOND(GravityModel):
    """Mesh (FDM/FFT) MOND: needs the reference's mesh solvers, which this
    direct-summation library does not provide."""


@dataclass
class GravityConfig:
    ForceSofteningTable: np.ndarray
    model: GravityModel = None
    MOND_nuIndex: float = 2.0


@dataclass
class DeviceConfig:
    type: DeviceType
    GPU_NumThreads: int
    GPU_Algorithm: GPUAlgorithm
    precision: str


@dataclass
class Constants:
    G: float
    ACC0: float = 1.2e-10    # MOND acceleration scale, in the caller's units


@dataclass
class OutputConfig:
    dir: str
    func: Optional[Callable]
    pot: bool
    SaveRestart: bool
    SaveRestartFreq: float
    format2: bool = True
    acc: bool = False


@dataclass
class SimConfig:
    time: TimeConfig
    grav: GravityConfig
    device: DeviceConfig
    constants: Constants
    output: OutputConfig
    floattype: type
    solver: DirectSum = field(default_factory=DirectSum)


@dataclass
class TimeInfo:
    """src/base/Config.jl:723-766"""
    dt: float = 0.0
    system_time_int: int = 0
    system_time_float: float = 0.0
    last_system_time_int: int = 0
    last_system_time_float: float = 0.0
    next_output_time_int: int = 0
    next_output_time_float: float = 0.0
    redshift: float = 0.0
    scalefactor: float = 1.0
    min_endstep: int = 0
    stepcount: int = 0


class StarData:
    """Column view of the particle data (the reference's StructArray of Star)."""

    def __init__(self, Pos, Vel=None, Mass=None, Collection=None, ID=None, Acc=None,
                 Potential=None, Ti_endstep=None, Ti_begstep=None):
        self.Pos = np.array(Pos, dtype=np.float64).reshape(-1, 3)
        n = self.Pos.shape[0]
        self.Vel = np.zeros((n, 3)) if Vel is None else np.array(Vel, dtype=np.float64).reshape(n, 3)
        self.Mass = np.zeros(n) if Mass is None else np.array(Mass, dtype=np.float64).reshape(n)
        self.Collection = (np.full(n, STAR, dtype=np.uint8) if Collection is None
                           else np.array(Collection, dtype=np.uint8).reshape(n))
        self.ID = np.arange(1, n + 1, dtype=np.int64) if ID is None else np.array(ID, dtype=np.int64)
        self.Acc = np.zeros((n, 3)) if Acc is None else np.array(Acc, dtype=np.float64).reshape(n, 3)
        self.Potential = np.zeros(n) if Potential is None else np.array(Potential, dtype=np.float64)
        self.Ti_endstep = np.zeros(n, dtype=np.int64) if Ti_endstep is None else np.array(Ti_endstep, dtype=np.int64)
        self.Ti_begstep = np.zeros(n, dtype=np.int64) if Ti_begstep is None else np.array(Ti_begstep, dtype=np.int64)

    def __len__(self):
        return self.Pos.shape[0]


_TIMERS = ("TOTAL", "FORCE", "DRIFT", "KICK", "ANALYSIS", "POSTSTEP", "OUTPUT", "PLOT")


class Simulation:
    """Simulation(data; kw...) for DirectSum on GPU (src/base/Config.jl:968)."""

    def __init__(self, data: StarData, *, pids=(1,), device: DeviceType = None,
                 GPU_NumThreads: int = 32, GPU_Algorithm: GPUAlgorithm = None,
                 floattype=np.float64, precision: Optional[str] = None,
                 TimeBegin: float = 0.0, TimeEnd: float = 0.1,
                 TimeBetweenSnapshots: float = 0.01, TimeStep: float = 0.0,
                 TimeBase: int = 1 << 30, ErrTolTimestep: float = 0.025,
                 MinStep: float = 1.0e-8, MaxStep: float = math.inf,
                 TimeIntegrationAlgorithm: TimeIntegration = None,
                 ForceSofteningTable=None, G: float = 1.0, constants: Constants = None,
                 model: GravityModel = None, MOND_nuIndex: float = 2.0,
                 OutputDir: Optional[str] = None, output_func: Optional[Callable] = None,
                 pot: bool = False, acc: bool = False, format2: bool = True,
                 SaveRestart: bool = False, SaveRestartFreq: float = 1.0,
                 analysers: Optional[Dict[str, Callable]] = None,
                 bgforce: Optional[list] = None, poststep: Optional[List[Callable]] = None,
                 write_timing_csv: bool = False):
        device = GPU() if device is None else device
        if isinstance(device, CPU):
            raise NotImplementedError(
                "this library serves the (DirectSum, GPU) methods only; the CPU path is "
                "the reference's own src/directsum")
        if not isinstance(data, StarData):
            raise TypeError("data must be a StarData")
        if np.any((data.Collection < 1) | (data.Collection > 6)):
            raise ValueError("Collection codes must be in 1..6")
        GPU_Algorithm = AllPairs() if GPU_Algorithm is None else GPU_Algorithm
        if precision is None:
            precision = "fp32" if np.dtype(floattype) == np.float32 else "split"
        if precision not in nat.MODES:
            raise ValueError(f"unknown precision {precision!r}; one of {sorted(nat.MODES)}")
        G = constants.G if constants is not None else G
        acc0 = constants.ACC0 if constants is not None else Constants(G).ACC0
        model = Newton() if model is None else model
        if isinstance(model, QUMOND):
            raise NotImplementedError("QUMOND is solved on a mesh; this library serves direct summation")
        table = (np.full(6, 1.0e-4) if ForceSofteningTable is None
                 else np.array(np.broadcast_to(np.asarray(ForceSofteningTable, dtype=np.float64), (6,))))
        # TimeConfig  src/base/Config.jl:408-449
        if TimeStep == 0:
            if TimeBase < (1 << 10):
                import warnings
                warnings.warn("TimeBase is too small!")
            interval = (TimeEnd - TimeBegin) / TimeBase
            step_cfg = AdaptiveTimestep(MinStep, MaxStep, int(TimeBase), interval, ErrTolTimestep)
            bsi = int(math.floor(TimeBetweenSnapshots / interval))
        else:
            step_cfg = ConstantTimestep(float(TimeStep))
            bsi = 0
        algo = Leapfrog() if TimeIntegrationAlgorithm is None else TimeIntegrationAlgorithm
        self.config = SimConfig(
            time=TimeConfig(float(TimeBegin), float(TimeEnd), float(TimeBetweenSnapshots), bsi,
                            step_cfg, algo),
            grav=GravityConfig(table, model, float(MOND_nuIndex)),
            device=DeviceConfig(device, int(GPU_NumThreads), GPU_Algorithm, precision),
            constants=Constants(float(G), float(acc0)),
            output=OutputConfig(os.path.join(os.getcwd(), "output") if OutputDir is None else OutputDir,
                                output_func, bool(pot), bool(SaveRestart),
                                float(SaveRestartFreq), bool(format2), bool(acc)),
            floattype=floattype,
        )
        self.pids = list(pids)
        self.id = 1
        self.timeinfo = TimeInfo(
            next_output_time_int=bsi,
            next_output_time_float=float(TimeBegin) + float(TimeBetweenSnapshots),
            system_time_float=float(TimeBegin))
        self.loginfo = {"analysers": dict(analysers or {}), "timers": {k: 0.0 for k in _TIMERS}}
        self.outputinfo = {"snapshotcount": 0}
        self._writer = None
        self._next_restart = math.inf
        self._resume_tail = False
        self.bgforce = list(bgforce or [])
        self.poststep = list(poststep or [])
        self.write_timing_csv = write_timing_csv
        self._host = data
        self.ctx = nat.Context(len(data), nat.MODES[precision])
        self.ctx.set_params(G, table)
        self.ctx.upload(data.Pos, data.Vel, data.Mass, data.Collection)
        for f in self.bgforce:
            self.ctx.add_background(*f)
        if isinstance(model, MOND1983Milgrom):
            self.ctx.set_mond(acc0, MOND_nuIndex)
        self._adaptive = None
        if isinstance(step_cfg, AdaptiveTimestep):
            from .adaptive import AdaptiveState
            self._adaptive = AdaptiveState(self)

    # convenience
    def __len__(self):
        return len(self._host)

    def close(self):
        try:
            if self._writer is not None:
                self._writer.close()
                self._writer = None
        finally:
            self.ctx.close()


# ---------------------------------------------------------------------------
# data access
# ---------------------------------------------------------------------------
def get_all_data(sim: Simulation) -> StarData:
    """D2H copy of the device state in the caller's order
    (src/base/Config.jl:1228-1230)."""
    d = sim.ctx.download(pos=True, vel=True, acc=True, pot=True)
    h = sim._host
    out = StarData(d["pos"], d["vel"], h.Mass, h.Collection, h.ID, d["acc"], d["pot"])
    if sim._adaptive is not None:
        out.Ti_endstep, out.Ti_begstep = sim._adaptive.download_steps()
    return out


get_local_data = get_all_data


def preprocessdata(sim: Simulation) -> None:
    """preprocessdata(sim, ::DirectSum, ::GPU)  (src/run.jl:9-11) sets the
    preferred units of the run.  This library holds unitless values (the
    caller strips units when it builds the StarData), so there is nothing to
    do; run(sim) needs no call to it."""
    return None


def numlocal(sim: Simulation) -> int:
    """length(get_local_data(sim))  (src/base/Config.jl:1203-1205)"""
    return len(sim)


def pack_pos(sim: Simulation):
    """x, y, z = pack_pos(sim)  (src/base/Plot.jl:1-5): the positions as three
    arrays in the caller's order (unitless, as the library holds them),
    downloaded without the other fields."""
    pos = sim.ctx.download(pos=True, vel=False, acc=False)["pos"]
    return pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()


def softlen(c, sim: Simulation) -> float:
    """ForceSofteningTable[Int(c)]  (src/tools/SofteningLength.jl:10-24)"""
    return float(sim.config.grav.ForceSofteningTable[int(c) - 1])


def set_softlen(sim: Simulation, *args):
    """set_softlen!(sim, h) / set_softlen!(sim, c, h)  (src/tools/SofteningLength.jl:64-93)"""
    if len(args) == 1:
        sim.config.grav.ForceSofteningTable[:] = float(args[0])
    elif len(args) == 2:
        sim.config.grav.ForceSofteningTable[int(args[0]) - 1] = float(args[1])
    else:
        raise TypeError("set_softlen(sim, h) or set_softlen(sim, collection, h)")
    sim.ctx.set_params(sim.config.constants.G, sim.config.grav.ForceSofteningTable)


set_softlen_ = set_softlen  # the reference's set_softlen!


# ---------------------------------------------------------------------------
# timers  (src/base/Timing.jl)
# ---------------------------------------------------------------------------
def add_timer(sim, name, t0, t1):
    sim.loginfo["timers"][name] += (t1 - t0) * 1e-9


# ---------------------------------------------------------------------------
# forces / potentials
# ---------------------------------------------------------------------------
def compute_force(sim: Simulation, pos=None, SoftLength=None):
    """compute_force(sim) and compute_force(sim, pos, h)."""
    if pos is not None:
        if SoftLength is None:
            raise TypeError("compute_force(sim, pos, SoftLength)")
        single = np.asarray(pos).ndim == 1
        acc = sim.ctx.force_at_points(np.atleast_2d(pos), SoftLength)
        return acc[0] if single else acc
    t0 = time.time_ns()
    if sim._adaptive is not None:
        sim._adaptive.compute_force()
    else:
        sim.ctx.compute_force()
    sim.ctx.sync()
    add_timer(sim, "FORCE", t0, time.time_ns())


def compute_potential(sim: Simulation, pos=None, SoftLength=None):
    """compute_potential(sim) (adds the reference's self term,
    src/base/Potential.jl:20-24) and compute_potential(sim, pos, h)."""
    if pos is not None:
        if SoftLength is None:
            raise TypeError("compute_potential(sim, pos, SoftLength)")
        single = np.asarray(pos).ndim == 1
        pot = sim.ctx.potential_at_points(np.atleast_2d(pos), SoftLength)
        return pot[0] if single else pot
    sim.ctx.compute_potential(self_term=True)


# ---------------------------------------------------------------------------
# energies / momenta  (src/base/Energy.jl)
# ---------------------------------------------------------------------------
def total_kinetic(sim: Simulation) -> float:
    return sim.ctx.energy(potential=False)["K"]


def total_potential(sim: Simulation) -> float:
    """0.5 * sum(m * phi), phi recomputed without the self term
    (src/directsum/potential.jl:150-153)."""
    return sim.ctx.energy(potential=True)["W"]


def total_energy(sim: Simulation) -> float:
    e = sim.ctx.energy(potential=True)
    return e["K"] + e["W"]


def total_momentum(sim: Simulation, axis: Optional[str] = None):
    P = sim.ctx.energy(potential=False)["P"]
    if axis is None:
        return P
    return float(P["xyz".index(axis)])


def total_angular_momentum(sim: Simulation):
    return sim.ctx.energy(potential=False)["L"]


# ---------------------------------------------------------------------------
# time integration  (src/base/Timestep.jl, src/directsumgpu/timestep.jl)
# ---------------------------------------------------------------------------
def init_timesteps(sim: Simulation):
    """src/directsumgpu/timestep.jl:32-47 (ConstantTimestep => Ti_endstep = 0)."""
    if sim._adaptive is not None:
        sim._adaptive.init_timesteps()


def leapfrog_half_kick(sim: Simulation, forward: int):
    """src/directsumgpu/timestep.jl:194-214 / :178-192."""
    t0 = time.time_ns()
    if sim._adaptive is not None:
        sim._adaptive.half_kick(forward)
    else:
        dt = sim.timeinfo.dt = sim.config.time.step.dt
        sim.ctx.half_kick(forward, dt)
    add_timer(sim, "KICK", t0, time.time_ns())


def drift_particles_const_timestep(sim: Simulation, next_float: float):
    """src/directsum/timestep.jl:23-40 + src/base/Timestep.jl:181-191: the drift
    amount is system_time_float - last_system_time_float (the CPU semantics;
    the reference GPU kernel's fixed-dt drift, src/directsumgpu/timestep.jl:67-74,
    double-drifts at snapshot times and is not reproduced)."""
    ti = sim.timeinfo
    ti.last_system_time_float = ti.system_time_float
    ti.system_time_float = next_float
    sim.ctx.drift(ti.system_time_float - ti.last_system_time_float)


def find_next_sync_point_and_drift(sim: Simulation):
    """src/base/Timestep.jl:197-215 (ConstantTimestep)."""
    if sim._adaptive is not None:
        return sim._adaptive.find_next_sync_point_and_drift()
    ti = sim.timeinfo
    next_system_time = ti.system_time_float + sim.config.time.step.dt
    while ti.next_output_time_float <= next_system_time:
        t0 = time.time_ns()
        drift_particles_const_timestep(sim, ti.next_output_time_float)
        add_timer(sim, "DRIFT", t0, time.time_ns())
        t0 = time.time_ns()
        _output(sim)
        add_timer(sim, "OUTPUT", t0, time.time_ns())
        ti.next_output_time_float += sim.config.time.BetweenSnapshots
    t0 = time.time_ns()
    drift_particles_const_timestep(sim, next_system_time)
    add_timer(sim, "DRIFT", t0, time.time_ns())


def advance_and_find_timestep(sim: Simulation):
    """src/directsumgpu/timestep.jl:164-176 (ConstantTimestep: Vel += Acc*dt)."""
    t0 = time.time_ns()
    if sim._adaptive is not None:
        sim._adaptive.advance_and_find_timestep()
    else:
        dt = sim.timeinfo.dt = sim.config.time.step.dt
        sim.ctx.kick(dt)
    add_timer(sim, "KICK", t0, time.time_ns())


def step(sim: Simulation):
    """step(sim, DirectSum, GPU)  src/directsumgpu/timestep.jl:216-224."""
    find_next_sync_point_and_drift(sim)
    compute_force(sim)
    advance_and_find_timestep(sim)


def output(sim: Simulation):
    """Default output function: one Gadget-2 snapshot
    OutputDir/snapshot_NNNN.gadget2 (src/base/Output.jl:36-63), written by a
    worker thread from an asynchronous device copy so that the run goes on
    meanwhile.  The header follows src/base/Output.jl:9-25."""
    from .snapshot import SnapshotWriter, header_for
    cfg = sim.config.output
    os.makedirs(cfg.dir, exist_ok=True)
    if cfg.pot:
        compute_potential(sim)
    if sim._writer is None:
        sim._writer = SnapshotWriter(sim.ctx)
    fname = os.path.join(cfg.dir, "snapshot_%04d.gadget2" % sim.outputinfo["snapshotcount"])
    h = sim._host
    header = header_for(h.Collection, sim.timeinfo.system_time_float, sim.timeinfo.redshift)
    sim._writer.submit(fname, header, h.ID, h.Mass, h.Collection, acc=cfg.acc, pot=cfg.pot,
                       format2=cfg.format2)
    sim.outputinfo["snapshotcount"] += 1


def _output(sim: Simulation):
    f = sim.config.output.func
    if f is not None:
        f(sim)
    else:
        output(sim)


def _fast_forward_steps(sim: Simulation) -> int:
    """How many upcoming steps contain no snapshot, analyser, poststep, restart
    or stop-file event, so that run() can hand them to nb_kdk_steps in one
    call.  Returns 0 when the next step needs the host."""
    if sim._adaptive is not None or sim.loginfo["analysers"] or sim.poststep:
        return 0
    if sim.config.output.SaveRestart or sim.write_timing_csv:
        return 0
    ti = sim.timeinfo
    dt = sim.config.time.step.dt
    t = ti.system_time_float
    k = 0
    while True:
        t_next = t + dt
        if ti.next_output_time_float <= t_next or t_next >= sim.config.time.End:
            return k
        t = t_next
        k += 1
        if k >= 256:   # the host checks the stop file between chunks
            return k


def run(sim: Simulation, use_graph: bool = True):
    """run(sim)  src/run.jl:46-184 for DirectSum on GPU.

    Stretches of steps with no host-side event are issued as one
    nb_kdk_steps call (identical step sequence, no host round trip)."""
    cfg = sim.config
    outdir = cfg.output.dir
    stopfile = os.path.join(outdir, "stop")
    os.makedirs(outdir, exist_ok=True)                     # run.jl:63
    if os.path.isfile(stopfile):
        os.remove(stopfile)
    compute_force(sim)                                       # run.jl:76
    init_timesteps(sim)                                      # run.jl:78
    if cfg.output.pot:
        compute_potential(sim)                               # run.jl:80-82
    if isinstance(cfg.time.algorithm, Leapfrog):
        leapfrog_half_kick(sim, +1)                          # run.jl:85-87
    _output(sim)                                             # run.jl:90
    sim._next_restart = cfg.output.SaveRestartFreq * (cfg.time.End - cfg.time.Begin)
    return _main_loop(sim, use_graph, append=False, run_loop=True)


def restart(sim: Simulation, use_graph: bool = True):
    """restart(sim)  src/run.jl:186-189: reopen the loggers for appending and
    continue the main loop of a Simulation returned by loadrestart.  The
    restart file is written inside a loop iteration (run.jl:138-141), so the
    rest of that iteration (end check, step counter) is finished first."""
    ti = sim.timeinfo
    run_loop = True
    if getattr(sim, "_resume_tail", False):
        sim._resume_tail = False
        run_loop = ti.system_time_float < sim.config.time.End
        ti.stepcount += 1
        ti.last_system_time_float = ti.system_time_float
    os.makedirs(sim.config.output.dir, exist_ok=True)
    return _main_loop(sim, use_graph, append=True, run_loop=run_loop)


def _main_loop(sim: Simulation, use_graph: bool, append: bool, run_loop: bool):
    """The while-loop and the closing half kick of src/run.jl:100-184."""
    cfg = sim.config
    ti = sim.timeinfo
    outdir = cfg.output.dir
    stopfile = os.path.join(outdir, "stop")
    analysis = []
    anames = list(sim.loginfo["analysers"])
    aio = tio = None
    mode = "a" if append else "w"
    if anames:                                               # Logging.jl:13-16
        path = os.path.join(outdir, "analysis.csv")
        fresh = not (append and os.path.isfile(path))
        aio = open(path, mode if not fresh else "w")
        if fresh:
            aio.write(",".join(["time"] + anames) + "\n")
    if sim.write_timing_csv:                                 # Logging.jl:8-11
        path = os.path.join(outdir, "timing.csv")
        fresh = not (append and os.path.isfile(path))
        tio = open(path, mode if not fresh else "w")
        if fresh:
            tio.write(",".join(["time"] + list(_TIMERS)) + "\n")
    restart_every = cfg.output.SaveRestartFreq * (cfg.time.End - cfg.time.Begin)
    state = run_loop
    try:
        while state:
            t_total = time.time_ns()
            # a stop file already present ends the run after one ordinary step,
            # as the per-step check of run.jl:153-157 would
            nfast = 0 if os.path.isfile(stopfile) else _fast_forward_steps(sim)
            if nfast >= 2:
                dt = cfg.time.step.dt
                t0 = time.time_ns()
                sim.ctx.kdk_steps(nfast, dt, ti.system_time_float, use_graph=use_graph)
                sim.ctx.sync()
                add_timer(sim, "FORCE", t0, time.time_ns())
                for _ in range(nfast):
                    ti.last_system_time_float = ti.system_time_float
                    ti.system_time_float = ti.system_time_float + dt
                ti.dt = dt
                ti.stepcount += nfast
                add_timer(sim, "TOTAL", t_total, time.time_ns())
                if os.path.isfile(stopfile):                 # run.jl:153-157, once per chunk
                    import warnings
                    warnings.warn("Interrupted by user")
                    state = False
                continue
            step(sim)                                        # run.jl:112
            t0 = time.time_ns()
            if anames:                                       # run.jl:115-119, Analyse.jl:7-22
                row = [ti.system_time_float] + [sim.loginfo["analysers"][k](sim) for k in anames]
                analysis.append(row)
                aio.write(",".join(_csv(v) for v in row) + "\n")
                aio.flush()
            add_timer(sim, "ANALYSIS", t0, time.time_ns())
            t0 = time.time_ns()
            for ps in sim.poststep:                          # run.jl:122-128
                ps(sim)
            add_timer(sim, "POSTSTEP", t0, time.time_ns())
            if cfg.output.SaveRestart and ti.system_time_float >= sim._next_restart:  # run.jl:138-141
                from .restart import saverestart
                sim._next_restart += restart_every
                saverestart(sim)
            if os.path.isfile(stopfile):                     # run.jl:153-157
                import warnings
                warnings.warn("Interrupted by user")
                state = False
            if ti.system_time_float >= cfg.time.End:        # run.jl:158-160
                state = False
            ti.stepcount += 1                                # run.jl:163-164
            ti.last_system_time_float = ti.system_time_float
            add_timer(sim, "TOTAL", t_total, time.time_ns())
            if tio is not None:                              # Timing.jl:32-37
                tio.write(",".join([repr(ti.system_time_float)] +
                                   [str(int(sim.loginfo["timers"][k])) for k in _TIMERS]) + "\n")
                tio.flush()
        if isinstance(cfg.time.algorithm, Leapfrog):
            leapfrog_half_kick(sim, -1)                      # run.jl:171-173
        sim.ctx.sync()
    finally:
        for f in (aio, tio):
            if f is not None:
                f.close()
    if sim._writer is not None:
        sim._writer.flush()                                  # every snapshot is on disk
    if anames:
        return np.array(analysis)
    return None


def _csv(v) -> str:
    """csvstring  src/base/Analyse.jl:1-5: vectors are quoted."""
    a = np.asarray(v)
    if a.ndim == 0:
        return repr(float(a)) if np.issubdtype(a.dtype, np.floating) else str(a.item())
    return '"' + "[" + ", ".join(repr(float(x)) for x in a.ravel()) + "]" + '"'
