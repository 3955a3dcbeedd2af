"""Adaptive (block, power-of-two) timesteps for the GPU Simulation: the host
side of the reference's AdaptiveTimestep path (src/base/Timestep.jl:1-245,
src/directsumgpu/timestep.jl:1-154), driving the nb_adaptive_* entry points.

The integer timeline: TimeInterval = (End - Begin) / TimeBase; particle i is
active when Ti_endstep[i] equals the system's integer time.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as nat
from ._native import call

_P = C.c_void_p
_I64 = C.POINTER(C.c_int64)

nat.register({
    "nb_adaptive_init": (C.c_int, [_P, C.c_double, C.c_int64, C.c_double, C.c_double, C.c_double]),
    "nb_adaptive_half_kick": (C.c_int, [_P, C.c_int32]),
    "nb_adaptive_min_endstep": (C.c_int, [_P, _I64]),
    "nb_adaptive_drift": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "nb_adaptive_compute_force": (C.c_int, [_P, C.c_int64, _I64]),
    "nb_adaptive_kick": (C.c_int, [_P, C.c_int64]),
    "nb_adaptive_get_steps": (C.c_int, [_P, _I64, _I64]),
    "nb_adaptive_set_steps": (C.c_int, [_P, _I64, _I64]),
})


class AdaptiveState:
    def __init__(self, sim):
        self.sim = sim
        self.cfg = sim.config.time.step
        self.initialised = False
        self.nactive_history = []

    # init_timesteps  src/base/Timestep.jl:29-44
    def init_timesteps(self):
        c = self.cfg
        call("nb_adaptive_init", self.sim.ctx.handle, float(c.TimeInterval), int(c.TimeBase),
             float(c.ErrTolTimestep), float(c.MinStep), float(c.MaxStep))
        self.initialised = True

    # compute_force with the active filter  src/directsum/gravity.jl:94-98
    def compute_force(self):
        if not self.initialised:
            # Ti_endstep starts at 0 == system_time_int: every particle is active
            self.sim.ctx.compute_force()
            return
        na = C.c_int64()
        call("nb_adaptive_compute_force", self.sim.ctx.handle,
             int(self.sim.timeinfo.system_time_int), C.byref(na))
        self.nactive_history.append(na.value)

    # leapfrog_half_kick (adaptive)  src/base/Timestep.jl:105-114
    def half_kick(self, forward: int):
        call("nb_adaptive_half_kick", self.sim.ctx.handle, int(forward))

    def _drift_to(self, next_int: int):
        """drift_particles_adaptive_timestep  src/directsum/timestep.jl:1-21"""
        ti = self.sim.timeinfo
        ti.last_system_time_int = ti.system_time_int
        ti.system_time_int = int(next_int)
        ti.system_time_float = self.sim.config.time.Begin + ti.system_time_int * self.cfg.TimeInterval
        ti.dt = (ti.system_time_int - ti.last_system_time_int) * self.cfg.TimeInterval
        call("nb_adaptive_drift", self.sim.ctx.handle, int(ti.last_system_time_int),
             int(ti.system_time_int))

    # find_next_sync_point_and_drift (adaptive)  src/base/Timestep.jl:224-245
    def find_next_sync_point_and_drift(self):
        from .simulation import _output, add_timer
        sim = self.sim
        ti = sim.timeinfo
        m = C.c_int64()
        call("nb_adaptive_min_endstep", sim.ctx.handle, C.byref(m))
        ti.min_endstep = m.value
        while ti.next_output_time_int <= ti.min_endstep:
            t0 = time.time_ns()
            self._drift_to(ti.next_output_time_int)
            add_timer(sim, "DRIFT", t0, time.time_ns())
            t0 = time.time_ns()
            _output(sim)
            add_timer(sim, "OUTPUT", t0, time.time_ns())
            ti.next_output_time_int += sim.config.time.BetweenSnapshotsInt
        t0 = time.time_ns()
        self._drift_to(ti.min_endstep)
        add_timer(sim, "DRIFT", t0, time.time_ns())

    # advance_and_find_timestep (adaptive)  src/base/Timestep.jl:47-82
    def advance_and_find_timestep(self):
        call("nb_adaptive_kick", self.sim.ctx.handle, int(self.sim.timeinfo.system_time_int))

    def set_steps(self, endstep, begstep):
        """Restore the integer timeline (caller order) after init_timesteps."""
        e = np.ascontiguousarray(endstep, dtype=np.int64)
        b = np.ascontiguousarray(begstep, dtype=np.int64)
        if e.shape != (len(self.sim),) or b.shape != e.shape:
            raise ValueError("endstep / begstep must have one entry per particle")
        call("nb_adaptive_set_steps", self.sim.ctx.handle, e.ctypes.data_as(_I64),
             b.ctypes.data_as(_I64))

    def download_steps(self):
        n = len(self.sim)
        e = np.zeros(n, dtype=np.int64)
        b = np.zeros(n, dtype=np.int64)
        if self.initialised:
            call("nb_adaptive_get_steps", self.sim.ctx.handle, e.ctypes.data_as(_I64),
                 b.ctypes.data_as(_I64))
        return e, b
