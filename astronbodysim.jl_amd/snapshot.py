"""Gadget-1/2 snapshot files and the asynchronous snapshot writer.

The reference writes snapshots through AstroIO's `write_gadget2` and reads
initial conditions with `read_gadget2` (src/base/Output.jl:9-50,
examples/04-collisions/04-collisions.jl:22).  AstroIO is not part of the
reference tree, so the file format here follows the published Gadget-2
layout (Springel 2005, Gadget-2 user guide section 6):

* every block is a Fortran unformatted record: int32 byte count, payload,
  the same int32 byte count;
* format 2 puts an 8-byte record in front of each block: a 4-character label
  and an int32 holding the next record's length plus 8;
* block order HEAD, POS, VEL, ID, MASS (only the types whose header mass is
  0), then for gas U, RHO, HSML, then optional POT and ACCE;
* particles are grouped by type 0..5 (gas, halo, disk, bulge, star,
  boundary), which are Collection codes 1..6 here;
* positions, velocities, masses, potentials and accelerations are float32,
  IDs uint32 (uint64 when an ID does not fit).

Snapshot writing overlaps the simulation: `SnapshotWriter.submit` enqueues a
device-to-host copy into a pinned slot (nb_snapshot_begin) and returns; a
worker thread waits for the copy (nb_snapshot_fetch) and writes the file.
"""
from __future__ import annotations

import os
import struct
import threading
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HEADER_FMT = "6i6d2d2i6I2i4d2i6Ii60s"
assert struct.calcsize("<" + _HEADER_FMT) == 256

LABELS = {"HEAD": b"HEAD", "POS": b"POS ", "VEL": b"VEL ", "ID": b"ID  ", "MASS": b"MASS",
          "U": b"U   ", "RHO": b"RHO ", "HSML": b"HSML", "POT": b"POT ", "ACCE": b"ACCE"}


@dataclass
class HeaderGadget2:
    """The 256-byte Gadget-2 header."""
    npart: List[int] = field(default_factory=lambda: [0] * 6)
    mass: List[float] = field(default_factory=lambda: [0.0] * 6)
    time: float = 0.0
    redshift: float = 0.0
    flag_sfr: int = 0
    flag_feedback: int = 0
    npartTotal: List[int] = field(default_factory=lambda: [0] * 6)
    flag_cooling: int = 0
    num_files: int = 1
    BoxSize: float = 0.0
    Omega0: float = 0.0
    OmegaLambda: float = 0.0
    HubbleParam: float = 0.0
    flag_stellarage: int = 0
    flag_metals: int = 0
    npartTotalHighWord: List[int] = field(default_factory=lambda: [0] * 6)
    flag_entropy_instead_u: int = 0
    fill: bytes = bytes(60)

    def pack(self, endian: str = "<") -> bytes:
        return struct.pack(endian + _HEADER_FMT, *self.npart, *self.mass, self.time,
                           self.redshift, self.flag_sfr, self.flag_feedback, *self.npartTotal,
                           self.flag_cooling, self.num_files, self.BoxSize, self.Omega0,
                           self.OmegaLambda, self.HubbleParam, self.flag_stellarage,
                           self.flag_metals, *self.npartTotalHighWord,
                           self.flag_entropy_instead_u, bytes(self.fill))

    @classmethod
    def unpack(cls, b: bytes, endian: str = "<") -> "HeaderGadget2":
        v = struct.unpack(endian + _HEADER_FMT, b)
        return cls(list(v[0:6]), list(v[6:12]), v[12], v[13], v[14], v[15], list(v[16:22]),
                   v[22], v[23], v[24], v[25], v[26], v[27], v[28], v[29], list(v[30:36]),
                   v[36], v[37])


def count_gadget_types(collection) -> List[int]:
    """Particles per Collection code 1..6 (src/base/Output.jl:1-7)."""
    c = np.asarray(collection, dtype=np.int64)
    return [int(np.count_nonzero(c == k)) for k in range(1, 7)]


def header_for(collection, time: float, redshift: float = 0.0) -> HeaderGadget2:
    """The header the reference writes for a simulation snapshot
    (src/base/Output.jl:9-25): counts per type, zero type masses (every mass
    goes to the MASS block), one file, Omega0 0.3, OmegaLambda 0.7, h 0.71."""
    counts = count_gadget_types(collection)
    return HeaderGadget2(npart=list(counts), mass=[0.0] * 6, time=float(time),
                         redshift=float(redshift), npartTotal=list(counts), num_files=1,
                         BoxSize=0.0, Omega0=0.3, OmegaLambda=0.7, HubbleParam=0.71)


# ---- writing ----------------------------------------------------------------
def _record(f, payload: bytes, endian: str):
    n = struct.pack(endian + "i", len(payload))
    f.write(n)
    f.write(payload)
    f.write(n)


def _block(f, label: str, payload: bytes, format2: bool, endian: str):
    if format2:
        _record(f, LABELS[label] + struct.pack(endian + "i", len(payload) + 8), endian)
    _record(f, payload, endian)


def write_gadget2(filename: str, header: HeaderGadget2, Pos, Vel, ID, Mass, Collection, *,
                  Acc=None, Potential=None, format2: bool = True, endian: str = "<"):
    """Write one Gadget-2 snapshot.  Particles are written grouped by type
    (stable within a type); `header.npart` must agree with `Collection`."""
    coll = np.asarray(Collection, dtype=np.int64)
    if list(header.npart) != count_gadget_types(coll):
        raise ValueError("header.npart does not match the particles' Collection codes")
    order = np.argsort(coll, kind="stable")
    ct = coll[order]
    f32 = np.dtype(endian + "f4")

    def vec(a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64)[order], dtype=f32).tobytes()

    ids = np.asarray(ID, dtype=np.int64)[order]
    if ids.size and (ids.min() < 0 or ids.max() >= 1 << 32):
        idb = ids.astype(endian + "u8").tobytes()
    else:
        idb = ids.astype(endian + "u4").tobytes()
    with open(filename, "wb") as f:
        _block(f, "HEAD", header.pack(endian), format2, endian)
        _block(f, "POS", vec(Pos), format2, endian)
        _block(f, "VEL", vec(Vel), format2, endian)
        _block(f, "ID", idb, format2, endian)
        # (vectorised: the writer runs beside the simulation's thread, and a
        # per-particle Python loop here would hold the interpreter lock)
        var = np.asarray(header.mass, dtype=np.float64)[ct - 1] == 0.0
        if var.any():
            m = np.asarray(Mass, dtype=np.float64)[order][var]
            _block(f, "MASS", m.astype(f32).tobytes(), format2, endian)
        ngas = header.npart[0]
        if ngas > 0:
            z = np.zeros(ngas, dtype=f32).tobytes()
            for lab in ("U", "RHO", "HSML"):
                _block(f, lab, z, format2, endian)
        if Potential is not None:
            _block(f, "POT", np.asarray(Potential, dtype=np.float64)[order].astype(f32).tobytes(),
                   format2, endian)
        if Acc is not None:
            _block(f, "ACCE", vec(Acc), format2, endian)


# ---- reading ----------------------------------------------------------------
def _detect(f):
    b = f.read(4)
    if len(b) < 4:
        raise ValueError("file too short for a Gadget snapshot")
    for endian in ("<", ">"):
        v = struct.unpack(endian + "i", b)[0]
        if v == 256:
            return endian, False
        if v == 8:
            return endian, True
    raise ValueError("not a Gadget-1/2 file: first record length is neither 256 nor 8")


def _read_record(f, endian: str, first: Optional[bytes] = None) -> bytes:
    b = first if first is not None else f.read(4)
    if len(b) < 4:
        raise EOFError
    n = struct.unpack(endian + "i", b)[0]
    payload = f.read(n)
    tail = f.read(4)
    if len(payload) != n or len(tail) != 4 or struct.unpack(endian + "i", tail)[0] != n:
        raise ValueError("corrupt Gadget record (length markers disagree)")
    return payload


def read_gadget2(filename: str):
    """Read a Gadget-1 (format 1) or Gadget-2 (format 2) file, either
    endianness.  Returns (header, blocks) where blocks maps "POS", "VEL",
    "ID", "Mass", "Collection" (and "POT"/"ACCE" when a format-2 file labels
    them) to arrays in file order, in the file's own units."""
    with open(filename, "rb") as f:
        endian, format2 = _detect(f)
        f.seek(0)
        labels: List[Optional[str]] = []
        records: List[bytes] = []
        while True:
            lab = None
            try:
                if format2:
                    hdr = _read_record(f, endian)
                    lab = hdr[:4].decode("ascii", "replace").strip()
                rec = _read_record(f, endian)
            except EOFError:
                break
            labels.append(lab)
            records.append(rec)
    if not records or len(records[0]) != 256:
        raise ValueError("missing 256-byte header block")
    header = HeaderGadget2.unpack(records[0], endian)
    n = int(sum(header.npart))
    ct = np.repeat(np.arange(1, 7, dtype=np.uint8), header.npart)
    f4 = np.dtype(endian + "f4")

    def get(k, name):
        if format2:
            if name not in labels:
                return None
            return records[labels.index(name)]
        return records[k] if k < len(records) else None

    out = {"Collection": ct}
    pos = get(1, "POS")
    vel = get(2, "VEL")
    idr = get(3, "ID")
    if pos is None or vel is None or idr is None:
        raise ValueError("snapshot lacks POS/VEL/ID blocks")
    out["POS"] = np.frombuffer(pos, dtype=f4).astype(np.float64).reshape(n, 3)
    out["VEL"] = np.frombuffer(vel, dtype=f4).astype(np.float64).reshape(n, 3)
    idt = np.dtype(endian + ("u8" if len(idr) == 8 * n else "u4"))
    out["ID"] = np.frombuffer(idr, dtype=idt).astype(np.int64)
    mass = np.asarray(header.mass, dtype=np.float64)[ct.astype(np.int64) - 1]
    var = mass == 0.0
    if var.any():
        mb = get(4, "MASS")
        if mb is None or len(mb) != 4 * int(var.sum()):
            raise ValueError("MASS block missing or of the wrong size")
        mass[var] = np.frombuffer(mb, dtype=f4)
    out["Mass"] = mass
    if format2:
        if "POT" in labels:
            out["POT"] = np.frombuffer(records[labels.index("POT")], dtype=f4).astype(np.float64)
        if "ACCE" in labels:
            out["ACCE"] = np.frombuffer(records[labels.index("ACCE")],
                                        dtype=f4).astype(np.float64).reshape(n, 3)
    return header, out


# ---- Gadget internal units --------------------------------------------------
def gadget_internal_G(G_cgs: float = 6.672e-8, UnitLength_in_cm: float = 3.085678e21,
                      UnitMass_in_g: float = 1.989e43,
                      UnitVelocity_in_cm_per_s: float = 1e5) -> float:
    """G in a Gadget parameter file's internal units (kpc, 1e10 Msun, km/s by
    default): G_cgs * UnitMass * UnitTime^2 / UnitLength^3 with
    UnitTime = UnitLength / UnitVelocity (Gadget-2 begrun.c set_units)."""
    unit_time = UnitLength_in_cm / UnitVelocity_in_cm_per_s
    return G_cgs * UnitMass_in_g * unit_time ** 2 / UnitLength_in_cm ** 3


# ---- asynchronous writer ----------------------------------------------------
class SnapshotWriter:
    """Writes snapshots on a worker thread while the simulation continues.

    `submit` takes the device copy in the calling thread (stream-ordered, so
    it sees exactly the state at the call) into one of two pinned slots and
    returns; the worker waits for the copy and writes the file.  A third
    snapshot waits for the first file to finish before reusing its slot.
    """

    def __init__(self, ctx):
        self.ctx = ctx
        self.pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="snapshot")
        self.jobs = [None, None]
        self.next_slot = 0
        self.lock = threading.Lock()

    def submit(self, filename: str, header: HeaderGadget2, ID, Mass, Collection, *,
               acc: bool = False, pot: bool = False, format2: bool = True):
        slot = self.next_slot
        self.next_slot ^= 1
        if self.jobs[slot] is not None:
            self.jobs[slot].result()          # slot free again (and errors surface here)
            self.jobs[slot] = None
        fields = ("pos", "vel") + (("acc",) if acc else ()) + (("pot",) if pot else ())
        self.ctx.snapshot_begin(slot, fields)
        ID = np.array(ID, copy=True)
        Mass = np.array(Mass, copy=True)
        Collection = np.array(Collection, copy=True)

        def job():
            d = self.ctx.snapshot_fetch(slot, fields)
            tmp = filename + ".part"
            write_gadget2(tmp, header, d["pos"], d["vel"], ID, Mass, Collection,
                          Acc=d.get("acc"), Potential=d.get("pot"), format2=format2)
            os.replace(tmp, filename)
            return filename

        self.jobs[slot] = self.pool.submit(job)
        return self.jobs[slot]

    def flush(self):
        """Wait for every submitted snapshot; re-raises a writer error."""
        for k in range(2):
            if self.jobs[k] is not None:
                job, self.jobs[k] = self.jobs[k], None
                job.result()

    def close(self):
        try:
            self.flush()
        finally:
            self.pool.shutdown(wait=True)
