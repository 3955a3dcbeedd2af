"""Batched ensembles: many independent systems of uneven size on one GPU
(nb_batch_* entry points of include/nbody_mi355x.h).

Each system follows the one-system semantics (pair law of
src/directsum/gravity.jl:8-28 with the system's own G and softening, constant
timestep KDK of src/directsumgpu/timestep.jl:156-224); the reference has no
ensemble object, so this module has no reference counterpart beyond that.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native as nat
from ._native import call, dptr


class BatchTuning(C.Structure):
    _fields_ = [
        ("targets_per_thread", C.c_int32),
        ("blocks", C.c_int32),
        ("steps_per_graph", C.c_int32),
        ("sort_interval", C.c_int32),
        ("resident", C.c_int32),
        ("far_path", C.c_int32),
        ("resident_waves", C.c_int32),
        ("reserved", C.c_int32),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)

nat.register({
    "nb_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, _I64, C.c_int32]),
    "nb_batch_destroy": (C.c_int, [_P]),
    "nb_batch_set_params": (C.c_int, [_P, _D, _D, _D]),
    "nb_batch_upload": (C.c_int, [_P, _D, _D, _D]),
    "nb_batch_download": (C.c_int, [_P, _D, _D, _D, _D]),
    "nb_batch_set_central": (C.c_int, [_P, C.c_int32]),
    "nb_batch_compute_force": (C.c_int, [_P]),
    "nb_batch_compute_potential": (C.c_int, [_P]),
    "nb_batch_half_kick": (C.c_int, [_P, C.c_int32]),
    "nb_batch_kdk_steps": (C.c_int, [_P, C.c_int64, C.c_int32]),
    "nb_batch_kdk_steps_tracked": (C.c_int, [_P, C.c_int64, C.c_int32, _D]),
    "nb_batch_energy": (C.c_int, [_P, _D, _D, _D, _D]),
    "nb_batch_sync": (C.c_int, [_P]),
    "nb_batch_set_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "nb_batch_get_tuning": (C.c_int, [_P, C.POINTER(BatchTuning)]),
    "nb_batch_autotune": (C.c_int, [_P]),
    "nb_batch_set_timing": (C.c_int, [_P, C.c_int32]),
    "nb_batch_get_timing": (C.c_int, [_P, _I64, _D]),
})


def _vec(v, n):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))


class Ensemble:
    """nsys independent systems; system s owns rows offsets[s]:offsets[s+1]."""

    def __init__(self, offsets, mode: str = "split"):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.nsys = off.size - 1
        self.offsets = off
        self.n = int(off[-1])
        h = _P()
        call("nb_batch_create", C.byref(h), self.nsys, off.ctypes.data_as(_I64), nat.MODES[mode])
        self._h = h
        self.mode = mode

    @property
    def handle(self):
        if self._h is None:
            raise nat.NBError("ensemble", nat.NB_ERR_STATE, "ensemble destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            nat.lib.nb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, G=1.0, soft=0.0, dt=1e-3):
        call("nb_batch_set_params", self.handle, dptr(_vec(G, self.nsys)),
             dptr(_vec(soft, self.nsys)), dptr(_vec(dt, self.nsys)))

    def upload(self, pos, vel, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(self.n)
        call("nb_batch_upload", self.handle, dptr(pos), dptr(vel), dptr(mass))

    def download(self, pos=True, vel=True, acc=True, pot=False):
        n = self.n
        bufs = {k: (np.empty(shape) if want else None) for k, want, shape in (
            ("pos", pos, (n, 3)), ("vel", vel, (n, 3)), ("acc", acc, (n, 3)), ("pot", pot, (n,)))}
        call("nb_batch_download", self.handle, dptr(bufs["pos"]), dptr(bufs["vel"]),
             dptr(bufs["acc"]), dptr(bufs["pot"]))
        return {k: v for k, v in bufs.items() if v is not None}

    def set_central(self, count: int):
        call("nb_batch_set_central", self.handle, int(count))

    def compute_force(self):
        call("nb_batch_compute_force", self.handle)

    def compute_potential(self):
        call("nb_batch_compute_potential", self.handle)

    def half_kick(self, forward: int):
        call("nb_batch_half_kick", self.handle, int(forward))

    def kdk_steps(self, nsteps: int, use_graph: bool = True):
        call("nb_batch_kdk_steps", self.handle, int(nsteps), 1 if use_graph else 0)

    def kdk_steps_tracked(self, nsteps: int, every: int):
        slots = nsteps // every
        out = np.zeros(max(slots, 1) * self.nsys)
        call("nb_batch_kdk_steps_tracked", self.handle, int(nsteps), int(every), dptr(out))
        return out[: slots * self.nsys].reshape(self.nsys, slots)

    def energy(self, potential: bool = True):
        K = np.zeros(self.nsys)
        W = np.zeros(self.nsys) if potential else None
        P = np.zeros((self.nsys, 3))
        L = np.zeros((self.nsys, 3))
        rc = nat.lib.nb_batch_energy(self.handle, dptr(K), dptr(W), dptr(P), dptr(L))
        if rc == nat.NB_ERR_NONFINITE:
            # NaN / Inf values are returned as they are, with a warning
            warnings.warn(nat.last_error(), RuntimeWarning, stacklevel=2)
        else:
            nat.check(rc, "nb_batch_energy")
        return {"K": K, "W": W, "P": P, "L": L}

    def sync(self):
        call("nb_batch_sync", self.handle)

    def set_tuning(self, targets_per_thread=1, blocks=0, steps_per_graph=16, sort_interval=16,
                   resident=1, far_path=0, resident_waves=0):
        t = BatchTuning(int(targets_per_thread), int(blocks), int(steps_per_graph),
                        int(sort_interval), int(resident), int(far_path), int(resident_waves))
        call("nb_batch_set_tuning", self.handle, C.byref(t))

    def get_tuning(self) -> dict:
        t = BatchTuning()
        call("nb_batch_get_tuning", self.handle, C.byref(t))
        return {"targets_per_thread": t.targets_per_thread, "blocks": t.blocks,
                "steps_per_graph": t.steps_per_graph, "sort_interval": t.sort_interval,
                "resident": t.resident, "far_path": t.far_path,
                "resident_waves": t.resident_waves}

    def autotune(self):
        call("nb_batch_autotune", self.handle)

    def set_timing(self, enable: bool):
        call("nb_batch_set_timing", self.handle, 1 if enable else 0)

    def get_timing(self):
        cnt = C.c_int64()
        ms = C.c_double()
        call("nb_batch_get_timing", self.handle, C.byref(cnt), C.byref(ms))
        return cnt.value, ms.value
