"""Host-side API over libaem: the drop-in for the reference's solver process.

The reference runs ``./FDTD inputs.txt`` as a child process
(``main_fdtd/master.py:438-441``) and reads results back from ``data.dat``.
:class:`Plan` and :class:`Simulation` expose the same steps in-process:

>>> plan = Plan.from_inputs("inputs.txt")        # parse + setup (no GPU needed)
>>> sim = Simulation(plan)                        # device allocation
>>> sim.run()                                     # all time steps on the GPU
>>> sim.finalize(); sim.write_outputs()           # data.dat + _geometry.bin
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import FIELDS, SERIES, check, lib


class Plan:
    """Parsed deck and setup products (aem_plan_* in include/aem.h)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self.view = _lib.PlanView.from_address(lib().aem_plan_view_get(handle))

    @staticmethod
    def options(**setup) -> Optional[_lib.PlanOptions]:
        """aem_plan_options holding the reference's setup constants with the
        given fields replaced (cpml_cells, cpml_grading, cpml_kappa_max,
        cpml_alpha_max, cpml_reflection, tfsf_inset, ntff_gap); None when
        nothing is replaced."""
        if not setup:
            return None
        o = _lib.PlanOptions()
        lib().aem_plan_options_init(C.byref(o))
        for k, val in setup.items():
            if k not in dict(_lib.PlanOptions._fields_):
                raise TypeError(f"unknown setup option {k!r}")
            setattr(o, k, val)
        return o

    @classmethod
    def from_inputs(cls, path: str, **setup) -> "Plan":
        """Parse a deck file; keyword arguments replace the reference's setup
        constants (see :meth:`options`)."""
        h = C.c_void_p()
        o = cls.options(**setup)
        check(lib().aem_plan_from_inputs_ex(os.fspath(path).encode(), C.byref(o) if o else None, C.byref(h)))
        return cls(h)

    @classmethod
    def from_text(cls, text: str, base_dir: str = ".", **setup) -> "Plan":
        h = C.c_void_p()
        o = cls.options(**setup)
        check(lib().aem_plan_from_text_ex(text.encode(), base_dir.encode(), C.byref(o) if o else None, C.byref(h)))
        return cls(h)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def view_ptr(self) -> int:
        return lib().aem_plan_view_get(self._h)

    @property
    def shape(self):
        """(NZ, NY, PX) of every field array (x fastest)."""
        v = self.view
        return (v.NZ, v.NY, v.PX)

    @property
    def time_steps(self) -> int:
        return self.view.time_steps

    def write_geometry(self, path: str) -> None:
        check(lib().aem_plan_write_geometry(self._h, os.fspath(path).encode()))

    def close(self) -> None:
        if self._h:
            lib().aem_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """One simulation resident on the current HIP device.

    ``ntff`` selects the far-field binning ("ordered": the reference's FP32
    sums bit for bit; "atomic" or "histogram": FP64 accumulation; "off": no
    time-domain far field, e.g. for objective-only runs), ``planes``
    the z planes each bulk thread marches, ``fuse_cpml=False`` forces the
    separate CPML kernels, ``single_pass=False`` separate H and E bulk
    launches, ``ntff_lists=False`` ordered binning by the row-scan kernel
    instead of the contributor lists (aem_sim_options in include/aem.h)."""

    def __init__(self, plan: Plan, ntff: str = "ordered", planes: int = 0, fuse_cpml: bool = True,
                 single_pass: bool = True, ntff_lists: bool = True):
        self.plan = plan
        opt = _lib.SimOptions(_lib.NTFF_BINNING[ntff], int(planes), 0 if fuse_cpml else 1, 0 if single_pass else 1,
                              0 if ntff_lists else 1)
        h = C.c_void_p()
        check(lib().aem_sim_create_ex(plan.handle, C.byref(opt), C.byref(h)))
        self._h = h

    @classmethod
    def from_inputs(cls, path: str) -> "Simulation":
        return cls(Plan.from_inputs(path))

    # -- stepping --------------------------------------------------------------
    def run(self, steps: Optional[int] = None, graph: bool = True) -> float:
        """Advance ``steps`` (default: all remaining) leapfrog steps; returns
        the device time per step in ms."""
        if steps is None:
            steps = self.plan.time_steps - self.steps_done
        check(lib().aem_run(self._h, int(steps), 1 if graph else 0))
        return float(lib().aem_last_run_ms(self._h))

    def time_phases(self, steps: int) -> np.ndarray:
        """Advance ``steps`` eagerly with HIP events between the step phases;
        returns ms per step for [H core bulk launch, rest of H, E core bulk
        launch, rest of E, outputs]."""
        ms = (C.c_float * 5)()
        check(lib().aem_time_phases(self._h, int(steps), ms, 5))
        return np.array(list(ms), np.float64)

    @property
    def core_cells(self) -> int:
        """Cells updated by the core bulk launch (phases 0 and 2 of
        :meth:`time_phases`)."""
        return int(lib().aem_core_cells(self._h))

    @property
    def single_pass(self) -> bool:
        """True if a step runs the core box as one single-pass launch (H and
        E together, timed as phase 0 of :meth:`time_phases`)."""
        return int(lib().aem_single_pass(self._h)) == 1

    @property
    def steps_done(self) -> int:
        return int(lib().aem_steps_done(self._h))

    def finalize(self) -> None:
        check(lib().aem_finalize(self._h))

    # -- SPICE co-simulation -------------------------------------------------------
    @property
    def spice_ports(self):
        """The deck's SPICE ports and circuit names (plan view): a dict with
        ``netlist``, ``ports`` (per port: dir, loc, size, bias, cap, node,
        source, capacitor) and ``extra_nodes``."""
        v = self.plan.view
        names = (v.spice_names or b"").decode().split("\n")
        ports = []
        for q in range(v.n_spice):
            sp = v.spice[q]
            node, src, cap = names[1 + 3 * q: 4 + 3 * q]
            ports.append(dict(dir=sp.dir, loc=tuple(sp.loc), size=tuple(sp.size), bias=sp.v_bias, cap=sp.cap,
                              node=node, source=src, capacitor=cap))
        extra = names[1 + 3 * v.n_spice: 1 + 3 * v.n_spice + v.n_spice_extra]
        return dict(netlist=names[0] if v.n_spice else "", ports=ports, extra_nodes=extra)

    def port_exchange(self, v_in=None, current: bool = True) -> Optional[np.ndarray]:
        """One exchange with the host circuit (aem_port_exchange): finish the
        step waiting for node voltages ``v_in`` (ports, then extra nodes; None
        when no step waits), then, if ``current``, run the next step up to its
        SPICE gap currents and return them."""
        v = self.plan.view
        vin = None if v_in is None else np.ascontiguousarray(v_in, np.float32).reshape(-1)
        if vin is not None and vin.size != v.n_spice + v.n_spice_extra:
            raise ValueError(f"v_in needs {v.n_spice + v.n_spice_extra} node voltages")
        out = np.zeros(max(v.n_spice, 1), np.float32)
        check(lib().aem_port_exchange(self._h, None if vin is None else vin.ctypes.data,
                                      out.ctypes.data if current else None))
        return out[:v.n_spice] if current else None

    # -- outputs -----------------------------------------------------------------
    def series(self, kind: str, idx: int = 0, n: Optional[int] = None, imag: bool = False) -> np.ndarray:
        """An output series; ``imag`` selects the imaginary part of a kmax
        simulation's complex series (AEM_IMAG)."""
        n = self.plan.time_steps if n is None else n
        out = np.zeros(n, np.float32)
        check(lib().aem_get_series(self._h, SERIES[kind] | (_lib.IMAG if imag else 0), idx, out.ctypes.data, n))
        return out

    def cseries(self, kind: str, idx: int = 0, n: Optional[int] = None) -> np.ndarray:
        """A kmax simulation's complex series (real and imaginary part) as complex64."""
        return (self.series(kind, idx, n) + 1j * self.series(kind, idx, n, imag=True)).astype(np.complex64)

    def dft(self, kind: str, freqs_hz, idx: int = 0) -> np.ndarray:
        """Complex spectrum of an output series at ``freqs_hz``, computed on
        the device in FP64 (numpy FFT sign and scaling; see aem_get_dft)."""
        f = np.ascontiguousarray(freqs_hz, np.float64).reshape(-1)
        out = np.zeros(2 * f.size, np.float64)
        check(lib().aem_get_dft(self._h, SERIES[kind], idx, f.ctypes.data, f.size, out.ctypes.data))
        return out[0::2] + 1j * out[1::2]

    # -- differentiation ---------------------------------------------------------
    def set_objective(self, theta_deg: float, phi_deg: float, freq_hz: float, pol: str = "theta") -> None:
        """Accumulate F = r E_pol(theta, phi, f) by the frequency-domain
        near-to-far-field transform during the run (aem_objective_set)."""
        check(lib().aem_objective_set(self._h, float(theta_deg), float(phi_deg), float(freq_hz),
                                      0 if pol == "theta" else 1))

    def set_farfield_dft(self, angles, f0_hz: float, df_hz: float, n_freq: int) -> None:
        """Accumulate r E_theta and r E_phi at the directions ``angles``
        [(theta, phi) in degrees] and frequencies f0 + q df by a running DFT
        of the far-field box currents (aem_ffdft_set); call before stepping."""
        ang = np.asarray(angles, np.float32).reshape(-1, 2)
        th = np.ascontiguousarray(ang[:, 0])
        ph = np.ascontiguousarray(ang[:, 1])
        check(lib().aem_ffdft_set(self._h, len(ang), th.ctypes.data, ph.ctypes.data, float(f0_hz), float(df_hz),
                                  int(n_freq)))
        self._ff_shape = (len(ang), int(n_freq))

    def farfield_dft(self) -> np.ndarray:
        """Complex (n_angles, n_freq, 2): r E_theta and r E_phi."""
        na, nf = self._ff_shape
        out = np.zeros(na * nf * 4, np.float64)
        check(lib().aem_ffdft_get(self._h, out.ctypes.data, out.size))
        o = out.reshape(na, nf, 2, 2)
        return o[..., 0] + 1j * o[..., 1]

    def objective(self) -> complex:
        out = np.zeros(2, np.float64)
        check(lib().aem_objective_get(self._h, out.ctypes.data))
        return complex(out[0], out[1])

    def record_design(self, lo, hi) -> None:
        """Record the design box of cells [lo, hi] (1-based, inclusive) for
        the adjoint gradient (aem_adjoint_record); call before stepping."""
        a = np.ascontiguousarray(lo, np.int32)
        b = np.ascontiguousarray(hi, np.int32)
        check(lib().aem_adjoint_record(self._h, a.ctypes.data, b.ctypes.data))

    def eps_gradient(self, n_materials: int) -> np.ndarray:
        """d|F|^2 / d eps_r of deck materials 1..n_materials by the discrete
        adjoint (aem_adjoint_eps_gradient); after the full run."""
        out = np.zeros(n_materials, np.float64)
        check(lib().aem_adjoint_eps_gradient(self._h, int(n_materials), out.ctypes.data))
        return out

    def design_box(self, lo, hi) -> None:
        """Give the cells [lo, hi] (1-based, inclusive) per-cell permittivity
        (aem_design_box); call before stepping.  Also records the box for the
        adjoint."""
        a = np.ascontiguousarray(lo, np.int32)
        b = np.ascontiguousarray(hi, np.int32)
        check(lib().aem_design_box(self._h, a.ctypes.data, b.ctypes.data))
        self._design_shape = tuple(int(b[q] - a[q] + 1) for q in (2, 1, 0))

    @property
    def design_shape(self):
        """(nz, ny, nx) of the design box."""
        return self._design_shape

    def set_design_eps(self, eps) -> None:
        """Relative permittivity of every design cell, shape ``design_shape``
        (or flat, x fastest)."""
        a = np.ascontiguousarray(eps, np.float32).reshape(-1)
        check(lib().aem_set_design_eps(self._h, a.ctypes.data, a.size))

    def gradient(self, dJ_dF: complex, n_materials: int = 0) -> np.ndarray:
        """Gradient of a general real objective J(F) of the far-field
        objective F (aem_adjoint): ``dJ_dF`` = dJ/dRe F + i dJ/dIm F at the
        run's F.  Per design cell (shape ``design_shape``) with a design box,
        else per deck material 1..n_materials."""
        g = np.array([complex(dJ_dF).real, complex(dJ_dF).imag], np.float64)
        shape = getattr(self, "_design_shape", None)
        n = int(np.prod(shape)) if shape else int(n_materials)
        out = np.zeros(n, np.float64)
        check(lib().aem_adjoint(self._h, g.ctypes.data, out.ctypes.data, n))
        return out.reshape(shape) if shape else out

    def checkpoint(self, every: int) -> None:
        """Keep the adjoint's design record for ``every`` steps and snapshots of
        the state every ``every`` steps; gradients then recompute each interval
        (aem_adjoint_checkpoint).  Call before record_design / design_box."""
        check(lib().aem_adjoint_checkpoint(self._h, int(every)))

    def series_gradient(self, dJ_dV=None, dJ_dI=None, n_materials: int = 0) -> np.ndarray:
        """Gradient of a general real objective J of the lumped-port series
        (aem_adjoint_series): ``dJ_dV``, ``dJ_dI`` of shape (n_ports,
        time_steps) hold dJ/dV and dJ/dI of every port and step (either may be
        None).  Per design cell with a design box, else per deck material
        1..n_materials."""
        v = self.plan.view
        T, npt = int(v.time_steps), int(v.n_ports)

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            if a.shape != (npt, T):
                raise ValueError(f"series derivatives must have shape {(npt, T)}, got {a.shape}")
            return a
        wv, wi = arr(dJ_dV), arr(dJ_dI)
        shape = getattr(self, "_design_shape", None)
        n = int(np.prod(shape)) if shape else int(n_materials)
        out = np.zeros(n, np.float64)
        check(lib().aem_adjoint_series(self._h, None if wv is None else wv.ctypes.data,
                                       None if wi is None else wi.ctypes.data, out.ctypes.data, n))
        return out.reshape(shape) if shape else out

    def port_gradient(self, port: int, freq_hz: float, dJ_dG: complex, n_materials: int = 0) -> np.ndarray:
        """Gradient of a real J(G) of one port's voltage spectrum
        G = dft('port_v', [freq_hz], port) given dJ_dG = dJ/dRe G + i dJ/dIm G:
        dJ/dV(step n+1) = Re(conj(dJ_dG) exp(-2 pi i f n dt))."""
        v = self.plan.view
        T = int(v.time_steps)
        n = np.arange(T, dtype=np.float64)
        p = float(freq_hz) * float(v.dt) * n
        p -= np.floor(p)
        w = np.zeros((int(v.n_ports), T), np.float64)
        w[port] = np.real(np.conj(complex(dJ_dG)) * np.exp(-2j * np.pi * p))
        return self.series_gradient(dJ_dV=w, n_materials=n_materials)

    def design_gradient(self) -> np.ndarray:
        """d|F|^2 / d eps_r of every design cell, shape ``design_shape``
        (aem_adjoint_design_gradient); after the full run."""
        out = np.zeros(int(np.prod(self._design_shape)), np.float64)
        check(lib().aem_adjoint_design_gradient(self._h, out.ctypes.data, out.size))
        return out.reshape(self._design_shape)

    def adjoint_step(self, steps: int = 1) -> None:
        check(lib().aem_adjoint_step(self._h, int(steps)))

    def adjoint_field(self, name: str) -> np.ndarray:
        v = self.plan.view
        out = np.zeros(v.ncell, np.float32)
        check(lib().aem_adjoint_get_field(self._h, FIELDS[name], out.ctypes.data, v.ncell))
        return out.reshape(self.plan.shape)

    def set_adjoint_field(self, name: str, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, np.float32).reshape(-1)
        check(lib().aem_adjoint_set_field(self._h, FIELDS[name], a.ctypes.data, a.size))

    def field(self, name: str, imag: bool = False) -> np.ndarray:
        """A field component in the common layout; ``imag`` selects the
        imaginary part of a kmax simulation's complex field (AEM_IMAG)."""
        v = self.plan.view
        out = np.zeros(v.ncell, np.float32)
        check(lib().aem_get_field(self._h, FIELDS[name] | (_lib.IMAG if imag else 0), out.ctypes.data, v.ncell))
        return out.reshape(self.plan.shape)

    def video(self, name: str, first_step: int = 1, n_steps: Optional[int] = None) -> np.ndarray:
        """Frames [first_step, first_step + n_steps) of the deck's video
        slice of one component, shape (n_steps, size_b, size_a)
        (aem_get_video)."""
        v = self.plan.view
        if n_steps is None:
            n_steps = v.time_steps - first_step + 1
        out = np.zeros(n_steps * v.vid_size1 * v.vid_size2, np.float32)
        check(lib().aem_get_video(self._h, FIELDS[name], int(first_step), int(n_steps), out.ctypes.data, out.size))
        return out.reshape(n_steps, v.vid_size2, v.vid_size1)

    def set_field(self, name: str, values: np.ndarray, imag: bool = False) -> None:
        v = self.plan.view
        a = np.ascontiguousarray(values, np.float32).reshape(-1)
        if a.size != v.ncell:
            raise ValueError(f"expected {v.ncell} values, got {a.size}")
        check(lib().aem_set_field(self._h, FIELDS[name] | (_lib.IMAG if imag else 0), a.ctypes.data, a.size))

    def ntff_raw(self, which: int) -> np.ndarray:
        v = self.plan.view
        n_ang = self._n_angles()
        L = self._len_ff()
        out = np.zeros(n_ang * L, np.float64)
        check(lib().aem_get_ntff_raw(self._h, which, out.ctypes.data, out.size))
        return out.reshape(n_ang, L)

    def ff_time(self) -> np.ndarray:
        n = self._n_angles()
        out = np.zeros(n, np.float32)
        check(lib().aem_get_ff_time(self._h, out.ctypes.data, n))
        return out

    def write_outputs(self, path: Optional[str] = None) -> None:
        check(lib().aem_write_outputs(self._h, None if path is None else os.fspath(path).encode()))

    def close(self) -> None:
        if self._h:
            lib().aem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _n_angles(self) -> int:
        return int(self.plan.view.n_angles)

    def _len_ff(self) -> int:
        return int(self.plan.view.len_ff)


class Batch:
    """Independent simulations advanced together (aem_batch_* in include/aem.h).

    Replaces the reference's one-process-per-member fan-out for populations
    and sweeps (optimization_scripts/GA.py:11-65).  Members keep their own
    state: after :meth:`run`, read, finalize and write each member as usual.
    The batch keeps references to its members so they outlive it."""

    def __init__(self, sims: Sequence[Simulation]):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        check(lib().aem_batch_create(arr, len(self.sims), C.byref(h)))
        self._h = h

    def run(self, steps: Optional[int] = None, graph: bool = True) -> float:
        """Advance every member ``steps`` steps (default: what the first
        member has left); returns the device time per batched step in ms."""
        if steps is None:
            s0 = self.sims[0]
            steps = s0.plan.time_steps - s0.steps_done
        check(lib().aem_batch_run(self._h, int(steps), 1 if graph else 0))
        return float(lib().aem_batch_last_run_ms(self._h))

    def close(self) -> None:
        if self._h:
            lib().aem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
