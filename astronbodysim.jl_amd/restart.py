"""Restart files: saverestart / loadrestart / restart for the (DirectSum, GPU)
Simulation.

The reference serialises the whole Simulation, device arrays included, with
JLD2 into OutputDir/restart.<process id>.jld2 (src/restart.jl:14-19) and
reads it back with loadrestart (src/restart.jl:127-136); its restart(sim)
calls a main loop that does not exist (src/run.jl:186-189).  Here a restart
file is a NumPy .npz archive (read back with allow_pickle=False, so loading
it executes nothing) holding:

* the particle state in the caller's order: Pos, Vel, Acc, Potential, Mass,
  Collection, ID and, for adaptive steps, Ti_endstep / Ti_begstep;
* a JSON document with the configuration, TimeInfo, the snapshot counter,
  the timers and the next restart time.

Python callables (output function, analysers, post-step hooks) cannot be
stored; pass them again to loadrestart.  In FP64 mode a resumed run repeats
the uninterrupted one bit for bit.
"""
from __future__ import annotations

import dataclasses
import io
import json
import math
import os
from typing import Callable, Dict, List, Optional

import numpy as np

from . import simulation as S

FORMAT = "astronbodysim-mi355x-restart"
VERSION = 1


def _restart_path(sim) -> str:
    return os.path.join(sim.config.output.dir, "restart.%d.npz" % sim.id)


def _num(x):
    """JSON-safe float (inf/nan as strings)."""
    if isinstance(x, float) and not math.isfinite(x):
        return repr(x)
    return x


def _unnum(x):
    if isinstance(x, str) and x in ("inf", "-inf", "nan"):
        return float(x)
    return x


def saverestart(sim, filename: Optional[str] = None) -> str:
    """saverestart(sim)  src/restart.jl:14-19,25-32.  Returns the file name."""
    cfg = sim.config
    d = S.get_all_data(sim)
    step = cfg.time.step
    if isinstance(step, S.AdaptiveTimestep):
        stepd = {"kind": "adaptive", "MinStep": _num(step.MinStep), "MaxStep": _num(step.MaxStep),
                 "TimeBase": int(step.TimeBase), "ErrTolTimestep": step.ErrTolTimestep}
    else:
        stepd = {"kind": "constant", "dt": step.dt}
    meta = {
        "format": FORMAT, "version": VERSION, "n": len(sim),
        "precision": cfg.device.precision,
        "floattype": np.dtype(cfg.floattype).name,
        "GPU_NumThreads": cfg.device.GPU_NumThreads,
        "GPU_Algorithm": type(cfg.device.GPU_Algorithm).__name__,
        "algorithm": type(cfg.time.algorithm).__name__,
        "time": {"Begin": cfg.time.Begin, "End": cfg.time.End,
                 "BetweenSnapshots": cfg.time.BetweenSnapshots, "step": stepd},
        "grav": {"ForceSofteningTable": [float(v) for v in cfg.grav.ForceSofteningTable],
                 "model": type(cfg.grav.model).__name__ if cfg.grav.model is not None else "Newton",
                 "MOND_nuIndex": cfg.grav.MOND_nuIndex},
        "constants": {"G": cfg.constants.G, "ACC0": cfg.constants.ACC0},
        "output": {"dir": cfg.output.dir, "pot": cfg.output.pot, "acc": cfg.output.acc,
                   "format2": cfg.output.format2, "SaveRestart": cfg.output.SaveRestart,
                   "SaveRestartFreq": cfg.output.SaveRestartFreq},
        "bgforce": [[int(k), [float(v) for v in p]] for k, p in sim.bgforce],
        "write_timing_csv": bool(sim.write_timing_csv),
        "timeinfo": {k: _num(v) for k, v in dataclasses.asdict(sim.timeinfo).items()},
        "outputinfo": dict(sim.outputinfo),
        "timers": dict(sim.loginfo["timers"]),
        "next_restart": _num(float(sim._next_restart)),
        "adaptive_initialised": bool(sim._adaptive is not None and sim._adaptive.initialised),
    }
    arrays = {"Pos": d.Pos, "Vel": d.Vel, "Acc": d.Acc, "Potential": d.Potential,
              "Mass": d.Mass, "Collection": d.Collection, "ID": d.ID,
              "Ti_endstep": d.Ti_endstep, "Ti_begstep": d.Ti_begstep,
              "meta": np.frombuffer(json.dumps(meta).encode("utf-8"), dtype=np.uint8)}
    path = filename or _restart_path(sim)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    buf = io.BytesIO()
    np.savez(buf, **arrays)
    tmp = path + ".part"
    with open(tmp, "wb") as f:
        f.write(buf.getvalue())
    os.replace(tmp, path)
    return path


_ALGOS = {"AllPairs": S.AllPairs, "Tiled": S.Tiled}
_INTEGRATORS = {"Leapfrog": S.Leapfrog, "Euler": S.Euler}
_MODELS = {"Newton": S.Newton, "MOND1983Milgrom": S.MOND1983Milgrom}


def loadrestart(filename: str, *, output_func: Optional[Callable] = None,
                analysers: Optional[Dict[str, Callable]] = None,
                poststep: Optional[List[Callable]] = None):
    """loadrestart(filename)  src/restart.jl:127-136: rebuild the Simulation
    (device state included) from a restart file; continue it with restart."""
    with np.load(filename, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode("utf-8"))
        arr = {k: np.array(z[k]) for k in z.files if k != "meta"}
    if meta.get("format") != FORMAT or meta.get("version") != VERSION:
        raise ValueError(f"{filename} is not a restart file of this library (version {VERSION})")
    t = meta["time"]
    st = t["step"]
    kw = {}
    if st["kind"] == "adaptive":
        kw.update(TimeStep=0.0, TimeBase=st["TimeBase"], MinStep=_unnum(st["MinStep"]),
                  MaxStep=_unnum(st["MaxStep"]), ErrTolTimestep=st["ErrTolTimestep"])
    else:
        kw.update(TimeStep=st["dt"])
    g = meta["grav"]
    out = meta["output"]
    data = S.StarData(arr["Pos"], Vel=arr["Vel"], Mass=arr["Mass"], Collection=arr["Collection"],
                      ID=arr["ID"])
    sim = S.Simulation(
        data, precision=meta["precision"], floattype=np.dtype(meta["floattype"]).type,
        GPU_NumThreads=meta["GPU_NumThreads"], GPU_Algorithm=_ALGOS[meta["GPU_Algorithm"]](),
        TimeIntegrationAlgorithm=_INTEGRATORS[meta["algorithm"]](),
        TimeBegin=t["Begin"], TimeEnd=t["End"], TimeBetweenSnapshots=t["BetweenSnapshots"],
        ForceSofteningTable=g["ForceSofteningTable"], model=_MODELS[g["model"]](),
        MOND_nuIndex=g["MOND_nuIndex"],
        constants=S.Constants(meta["constants"]["G"], meta["constants"]["ACC0"]),
        OutputDir=out["dir"], output_func=output_func, pot=out["pot"], acc=out["acc"],
        format2=out["format2"], SaveRestart=out["SaveRestart"],
        SaveRestartFreq=out["SaveRestartFreq"], analysers=analysers,
        bgforce=[(k, p) for k, p in meta["bgforce"]], poststep=poststep,
        write_timing_csv=meta["write_timing_csv"], **kw)
    sim.ctx.upload_acc(arr["Acc"], arr["Potential"])
    for k, v in meta["timeinfo"].items():
        setattr(sim.timeinfo, k, _unnum(v))
    sim.outputinfo.update(meta["outputinfo"])
    sim.loginfo["timers"].update(meta["timers"])
    sim._next_restart = _unnum(meta["next_restart"])
    if sim._adaptive is not None and meta["adaptive_initialised"]:
        sim._adaptive.init_timesteps()
        sim._adaptive.set_steps(arr["Ti_endstep"], arr["Ti_begstep"])
    sim._resume_tail = True
    return sim
