"""Writers for the reference's positional ``inputs.txt`` deck.

The grammar is the read sequence of ``main_fdtd/fdtd_solver.f90:726-1298``
(non-kmax, non-SPICE build) as documented line by line in
``main_fdtd/master.py:51-295``.  :class:`Deck` holds the deck as data and
:meth:`Deck.text` renders it; the workload builders below produce the decks
used by the tests and the benchmark.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np


def _f(x) -> str:
    """Render a number the way a Fortran list-directed read accepts it."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    r = repr(float(x))
    return r.replace("e", "E")


def _row(*xs) -> str:
    return ",".join(_f(x) for x in xs)


@dataclass
class Material:
    """``volume`` material (f90:929-947): simple or Drude (``plasma``)."""
    mat_id: int
    eps: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    sigma: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    poles: Sequence[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = ()

    def lines(self) -> List[str]:
        out = ["volume", "plasma" if self.poles else "simple", _f(self.mat_id), _row(*self.eps), _row(*self.sigma)]
        if self.poles:
            out.append(_f(len(self.poles)))
            for wp, gam in self.poles:
                out += [_row(*wp), _row(*gam)]
        return out


@dataclass
class SheetMaterial:
    """``sheet`` material (f90:948-972): surface resistance in ohm/sq (0 = PEC)."""
    mat_id: int
    ohms_per_sq: float

    def lines(self) -> List[str]:
        return ["sheet", _f(self.mat_id), _f(self.ohms_per_sq)]


@dataclass
class Block:
    mat_id: int
    start: Tuple[int, int, int]
    size: Tuple[int, int, int]

    def lines(self) -> List[str]:
        return ["block", _f(self.mat_id), _row(*self.start), _row(*self.size)]


@dataclass
class Sphere:
    mat_id: int
    centre: Tuple[int, int, int]
    radius: int

    def lines(self) -> List[str]:
        return ["sphere", _f(self.mat_id), _row(*self.centre), _f(self.radius)]


@dataclass
class Cylinder:
    mat_id: int
    axis: str
    start: Tuple[int, int, int]
    length: int
    radius: int

    def lines(self) -> List[str]:
        return ["cylinder", _f(self.mat_id), self.axis, _row(*self.start), _row(self.length, self.radius)]


@dataclass
class Sheet:
    axis: str
    mat_id: int
    position: int
    start: Tuple[int, int]
    size: Tuple[int, int]

    def lines(self) -> List[str]:
        return [self.axis, _f(self.mat_id), _f(self.position), _row(*self.start), _row(*self.size)]


@dataclass
class Port:
    """Basic lumped port (f90:1125-1157); size as typed in master.py."""
    axis: str
    impedance: float
    start: Tuple[int, int, int]
    size: Tuple[int, int, int]

    def lines(self) -> List[str]:
        return ["basic", self.axis, _f(self.impedance), _row(*self.start), _row(*self.size)]


@dataclass
class SpicePort:
    """Basic SPICE port of the SPICE build's deck section (f90:1195-1286):
    gap axis, first cell, size as typed, DC bias and the circuit's names for
    the port node, its current source and its capacitor."""
    axis: str
    start: Tuple[int, int, int]
    size: Tuple[int, int, int] = (1, 1, 1)
    bias: float = 0.0
    node: str = "out"
    source: str = "I"
    capacitor: str = "Cp"

    def lines(self) -> List[str]:
        return ["basic", self.axis, _row(*self.start), _row(*self.size), _f(self.bias), self.node, self.source,
                self.capacitor]


@dataclass
class Deck:
    cells: Tuple[int, int, int]
    cell_size: Tuple[float, float, float]
    time_steps: int
    factor: float = 1.0
    f_center: float = 5e9
    pulse_type: int = 1
    boundaries: Tuple[int, int, int, int, int, int] = (0, 0, 0, 0, 0, 0)
    excitation: str = "antenna"                      # 'antenna' | 'plane wave' | 'none'
    amplitude: float = 1.0
    delay: float = 0.0
    angles: Tuple[float, float, float] = (0.0, 0.0, 0.0)   # plane wave theta, phi, pol (deg)
    excitation_port: int = 1
    phase_centre: Tuple[int, int, int] = (1, 1, 1)
    far_field: Sequence[Tuple[float, float]] = ()
    video: Optional[Tuple[str, int]] = None
    export_all: bool = False                         # per-step full-field files (f90:913)
    filename: str = "data.dat"
    materials: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    sheets: list = field(default_factory=list)
    ports: list = field(default_factory=list)
    optional_geometry: Optional[str] = None
    spice_ports: list = field(default_factory=list)
    netlist: str = "fdtd_netlist.cir"
    spice_extra_nodes: list = field(default_factory=list)
    igp: Optional[Tuple[str, int]] = None            # infinite ground plane: ('-z', height) etc. (f90:746-765)

    def text(self) -> str:
        L = [_row(*self.cells), _row(*self.cell_size), _f(self.time_steps), _f(self.factor), _f(self.f_center),
             _f(self.pulse_type), _row(*self.boundaries)]
        L += ["IGP", self.igp[0], _f(self.igp[1])] if self.igp else ["none"]
        L.append(self.excitation)
        if self.excitation == "plane wave":
            L += [_row(self.amplitude, self.delay), _row(*self.angles)]
        elif self.excitation == "antenna":
            L += [_row(self.amplitude, self.delay), _f(self.excitation_port)]
        L.append(_f(len(self.far_field)))
        if self.far_field:
            L.append(_row(*self.phase_centre))
            L += [_row(t, p) for t, p in self.far_field]
        if self.video:
            L += ["yes", self.video[0], _f(self.video[1])]
        else:
            L.append("no")
        L += ["yes" if self.export_all else "no", self.filename, _f(len(self.materials))]
        for m in self.materials:
            L += m.lines()
        if self.optional_geometry:
            L += ["yes", self.optional_geometry]
        else:
            L.append("no")
        L.append(_f(len(self.objects)))
        for o in self.objects:
            L += o.lines()
        L.append(_f(len(self.sheets)))
        for s in self.sheets:
            L += s.lines()
        L.append(_f(len(self.ports)))
        for p in self.ports:
            L += p.lines()
        L.append(_f(len(self.spice_ports)))
        if self.spice_ports:
            L.append(self.netlist)
            for p in self.spice_ports:
                L += p.lines()
            L.append(_f(len(self.spice_extra_nodes)))
            L += list(self.spice_extra_nodes)
        return "\n".join(L) + "\n"


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def gauss_f_center(spread_steps: float, dt: float) -> float:
    """f_center giving a Gaussian spread of ``spread_steps`` time steps
    (spread = 1/(2 pi f_center), f90:1712)."""
    return 1.0 / (2.0 * np.pi * spread_steps * dt)


def courant_dt(d: float, factor: float) -> float:
    """Time step of a cubic grid, FP64 value of f90:1312."""
    return 0.98 / (299792458.0 * np.sqrt(3.0 / d ** 2)) / factor


def dipole(n: int = 100, steps: int = 2000, d: float = 1e-3, angles: int = 37, arm: int = 20,
           factor: float = 0.98 / 0.99) -> Deck:
    """Radiating z-dipole of two sigma=1e8 wires with a 1-cell gap and a 50 ohm
    port at the centre; Gaussian of spread 40 dt and t_spread 160 dt.

    The reference fixes t_spread = 5*spread + delay (f90:1713), so a delay of
    -40 dt gives the 160 dt of the workload definition."""
    c = n // 2
    dt = courant_dt(d, factor)
    ff = [(float(t), 0.0) for t in np.linspace(0.0, 180.0, angles)] if angles else []
    return Deck(cells=(n, n, n), cell_size=(d, d, d), time_steps=steps, factor=factor,
                f_center=gauss_f_center(40.0, dt), pulse_type=1, excitation="antenna",
                amplitude=1.0, delay=-40.0 * dt, excitation_port=1, phase_centre=(c, c, c), far_field=ff,
                materials=[Material(4, sigma=(1e8, 1e8, 1e8))],
                objects=[Block(4, (c, c, c - arm), (1, 1, arm)), Block(4, (c, c, c + 1), (1, 1, arm))],
                ports=[Port("z", 50.0, (c, c, c), (0, 0, 1))])


def roofline_deck(n: int = 400, steps: int = 1100, angles: int = 0) -> Deck:
    """The 400^3 throughput workload: vacuum, 10-cell CPML, z-directed 50 ohm
    port at the centre, Gaussian spread 40 dt / t_spread 160 dt."""
    c = n // 2
    dt = courant_dt(1e-3, 0.98 / 0.99)
    ff = [(float(t), 0.0) for t in np.linspace(0.0, 180.0, angles)] if angles else []
    return Deck(cells=(n, n, n), cell_size=(1e-3, 1e-3, 1e-3), time_steps=steps, factor=0.98 / 0.99,
                f_center=gauss_f_center(40.0, dt), pulse_type=1, excitation="antenna", amplitude=1.0,
                delay=-40.0 * dt, excitation_port=1, phase_centre=(c, c, c), far_field=ff,
                ports=[Port("z", 50.0, (c, c, c), (0, 0, 1))])


def pec_sphere(n: int = 200, steps: int = 4000, radius: int = 40, d: float = 2e-3, angles: Sequence = None) -> Deck:
    """TF/SF scattering from a sigma=1e8 sphere: +z propagation, x polarised,
    differentiated Gaussian of spread 30 dt (pulse_type 2)."""
    c = n // 2
    dt = courant_dt(d, 1.0)
    ff = list(angles) if angles is not None else [(180.0, 0.0), (90.0, 0.0), (0.0, 0.0)]
    return Deck(cells=(n, n, n), cell_size=(d, d, d), time_steps=steps, factor=1.0,
                f_center=gauss_f_center(30.0, dt), pulse_type=2, excitation="plane wave", amplitude=1.0,
                angles=(0.0, 0.0, 270.0), phase_centre=(c, c, c), far_field=ff,
                materials=[Material(6, sigma=(1e8, 1e8, 1e8))], objects=[Sphere(6, (c, c, c), radius)])


def layered_slab(nz: int = 1200, nxy: int = 32, steps: int = 12000, d: float = 0.5e-3, seed_layers: int = 7,
                 seed_drude: int = 11) -> Deck:
    """Normal-incidence TF/SF through five anisotropic lossy layers, a 6-pole
    Drude slab and a 100 ohm/sq resistive sheet; periodic in x and y."""
    rng = np.random.default_rng(seed_layers)
    rng_d = np.random.default_rng(seed_drude)
    mats, objs = [], []
    z = 200
    for q in range(5):
        th = int(rng.integers(20, 81))
        eps = tuple(float(v) for v in rng.uniform(1.0, 6.0, 3))
        sig = tuple(float(v) for v in rng.uniform(0.0, 0.05, 3))
        mats.append(Material(10 + q, eps, sig))
        objs.append(Block(10 + q, (1, 1, z), (nxy, nxy, th)))
        z += th + 10
    poles = []
    for _ in range(6):
        wp = float(rng_d.uniform(1e10, 1e11))
        g = float(rng_d.uniform(1e8, 1e10))
        poles.append(((wp, wp, wp), (g, g, g)))
    mats.append(Material(20, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), poles))
    objs.append(Block(20, (1, 1, z), (nxy, nxy, 100)))
    z += 110
    dt = courant_dt(d, 1.0)
    return Deck(cells=(nxy, nxy, nz), cell_size=(d, d, d), time_steps=steps, factor=1.0,
                f_center=gauss_f_center(30.0, dt), pulse_type=2, boundaries=(1, 1, 1, 1, 0, 0),
                excitation="plane wave", amplitude=1.0, angles=(0.0, 0.0, 270.0),
                materials=mats + [SheetMaterial(30, 100.0)], objects=objs,
                sheets=[Sheet("z", 30, min(z, nz - 40), (1, 1), (nxy, nxy))])


def adjoint_deck(n: int = 160, steps: int = 4000, box: Tuple[int, int, int] = (40, 40, 20),
                 angle: Tuple[float, float] = (30.0, 0.0)) -> Tuple[Deck, Tuple[int, int, int], Tuple[int, int, int]]:
    """The differentiable workload: vacuum, 10-cell CPML, +z x-polarised TF/SF
    differentiated Gaussian centred on 10 GHz, one far-field angle, and a
    design box of ``box`` cells at the centre.  Returns (deck, lo, hi) of the
    box (1-based, inclusive)."""
    c = n // 2
    lo = tuple(c - b // 2 + 1 for b in box)
    hi = tuple(lo[a] + box[a] - 1 for a in range(3))
    d = Deck(cells=(n, n, n), cell_size=(1e-3, 1e-3, 1e-3), time_steps=steps, factor=1.0, f_center=10e9,
             pulse_type=2, excitation="plane wave", amplitude=1.0, angles=(0.0, 0.0, 270.0), phase_centre=(c, c, c),
             far_field=[angle])
    return d, lo, hi


def design_permittivity(shape: Tuple[int, int, int], seed: int = 5, lo: float = 1.0, hi: float = 4.0) -> np.ndarray:
    """Per-voxel eps_r uniform in [lo, hi] (PCG64 ``seed``), shape (nz, ny, nx)."""
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


def population(count: int = 64, seed: int = 3, steps: int = 3000, edges=range(80, 161, 16)) -> List[Deck]:
    """Independent dipole antennas of random size, arm and dielectric blocks."""
    rng = np.random.default_rng(seed)
    decks = []
    edges = list(edges)
    for _ in range(count):
        n = int(rng.choice(edges))
        arm = int(rng.integers(10, 31))
        arm = min(arm, n // 2 - 14)
        dk = dipole(n=n, steps=steps, angles=19, arm=arm)
        nb = int(rng.integers(0, 5))
        for b in range(nb):
            # blocks fit between the 13-cell margins (only bites below 46 cells)
            sz = [min(int(v), n - 26) for v in rng.integers(4, 21, 3)]
            st = [int(rng.integers(13, n - 12 - s)) for s in sz]
            dk.materials.append(Material(40 + b, (float(rng.uniform(1.0, 10.0)),) * 3, (0.0, 0.0, 0.0)))
            dk.objects.insert(0, Block(40 + b, tuple(st), tuple(sz)))
        decks.append(dk)
    return decks
