"""Reader for the solver output file ``data.dat``.

Layout of the reference writer (``main_fdtd/fdtd_solver.f90:8536-8645``): a
17-line header of ``name = value`` records, then labelled sections (incident
wave, periodic reflection/transmission, far field per angle, port voltages,
far-field time corrections, phase-centre incident).  The positions match what
``utility_scripts/post_processor.py:43-160`` consumes.

The kmax build's file (f90:8648-8785) has the same header, then the k vector
and mode records, complex series (real and imaginary part per line) and the
TE/TM reflected and transmitted waves in place of the periodic R/T; its
series are returned as complex64 arrays and ``kmax`` is set.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np


@dataclass
class DatFile:
    header: Dict[str, str] = field(default_factory=dict)
    time_steps: int = 0
    dt: float = 0.0
    incident: np.ndarray = None
    reflected: np.ndarray = None
    transmitted: np.ndarray = None
    e_theta: List[np.ndarray] = field(default_factory=list)
    e_phi: List[np.ndarray] = field(default_factory=list)
    angles: List[tuple] = field(default_factory=list)
    port_v: List[np.ndarray] = field(default_factory=list)
    port_z: List[float] = field(default_factory=list)
    spice_v: List[np.ndarray] = field(default_factory=list)   # SPICE build: per SPICE port / extra node
    ff_time: np.ndarray = None
    ff_pc: np.ndarray = None
    kmax: bool = False
    k_vec: tuple = None            # (k1, k2) of the two periodic axes
    mode_type: int = -1            # 0 TE, 1 TM
    te_refl: np.ndarray = None     # kmax TE/TM waves, as labelled by the writer (f90:8690-8730)
    tm_refl: np.ndarray = None
    te_trans: np.ndarray = None
    tm_trans: np.ndarray = None


def read_dat(path: str) -> DatFile:
    lines = [l.strip() for l in open(path).read().split("\n")]
    out = DatFile()
    for l in lines[:17]:
        if "=" in l:
            k, v = l.split("=", 1)
            out.header[k.strip()] = v.strip()
        elif l.startswith("simulation type"):
            out.header["simulation type"] = l[len("simulation type"):].strip()
    out.time_steps = int(out.header["number of time steps"])
    out.dt = float(out.header["time step"])
    secs = []
    for l in lines[17:]:
        if not l:
            continue
        try:
            vals = [float(q) for q in l.split()]
        except ValueError:
            secs.append((l, []))
            continue
        secs[-1][1].append(vals)
    T = out.time_steps
    out.kmax = any(name.startswith("simulation is k vector") for name, _ in secs)
    if out.kmax:
        def cplx(rows):
            return np.array([complex(r[0], r[1]) for r in rows], np.complex64)
        for name, rows in secs:
            if name.startswith("simulation is k vector"):
                out.k_vec = tuple(rows[0]) if rows else None
            elif name.startswith("mode type is"):
                out.mode_type = int(name.split(":")[-1])
            elif name.startswith("incident voltage") or name.startswith("incident E-field"):
                out.incident = cplx(rows)
            elif name.startswith("TE Refl"):
                out.te_refl = cplx(rows)
            elif name.startswith("TM Refl"):
                out.tm_refl = cplx(rows)
            elif name.startswith("TE Trans"):
                out.te_trans = cplx(rows)
            elif name.startswith("TM Trans"):
                out.tm_trans = cplx(rows)
            elif name.startswith("recieved voltage"):
                out.port_z.append(rows[0][0])
                out.port_v.append(cplx(rows[1:T + 1]))
        return out
    for name, rows in secs:
        if name.startswith("incident voltage") or name.startswith("incident E-field"):
            out.incident = np.array([r[0] for r in rows], np.float32)
        elif name.startswith("reflected wave"):
            out.reflected = np.array([r[0] for r in rows], np.float32)
        elif name.startswith("transmitted wave"):
            out.transmitted = np.array([r[0] for r in rows], np.float32)
        elif name.startswith("far field rE theta"):
            out.angles.append(tuple(rows[0]))
            out.e_theta.append(np.array([r[0] for r in rows[1:T + 1]], np.float32))
        elif name.startswith("far field rE phi"):
            out.e_phi.append(np.array([r[0] for r in rows[1:T + 1]], np.float32))
        elif name.startswith("recieved voltage at spice port"):
            out.spice_v.append(np.array([r[0] for r in rows[1:T + 1]], np.float32))
        elif name.startswith("recieved voltage"):
            out.port_z.append(rows[0][0])
            out.port_v.append(np.array([r[0] for r in rows[1:T + 1]], np.float32))
        elif name.startswith("far field phase correction"):
            out.ff_time = np.array([r[0] for r in rows], np.float32)
        elif name.startswith("incident for far field"):
            out.ff_pc = np.array([r[0] for r in rows], np.float32)
    return out


def read_fortran_record_file(path: str) -> np.ndarray:
    """Payload of a one-record Fortran unformatted sequential file (the video
    writers, f90:8788-8807): 4-byte length marker, float32 data, marker."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    return np.frombuffer(raw[4:4 + n], np.float32).copy()
