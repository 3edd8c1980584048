"""Host side of SPICE co-simulation over aem_port_exchange.

The reference's SPICE build drives ngspice from the solver loop
(main_fdtd/fdtd_solver.f90:5435-5665 and 6285-6520, main_fdtd/circuit.F90:
140-174): after the H phase of step n it alters each port's current source to
a one-point PWL holding the gap current I_n, resumes the transient analysis
to t = n dt (a "stop when time" breakpoint at every step, circuit.F90:
158-170), reads each node voltage and hard-sets the gap field from it.

:func:`run_cosim` is that loop over any host circuit object with a
``step(n, currents) -> node voltages`` method.  ngspice itself is not part of
this package (a caller binds libngspice and issues the same ``alter`` /
``resume`` commands inside ``step``); :class:`SeriesSourceCircuit` is an
exactly integrated linear circuit of the reference example's topology, used
by the tests and usable where that topology is all a deck needs.
"""
from __future__ import annotations

import math
from typing import Callable, Optional, Sequence

import numpy as np


class SeriesSourceCircuit:
    """Per SPICE port the netlist of the reference's SPICE example
    (examples/radiating_dipole_antenna_w_SPICE_source/master.py:325-337)::

        I   0   out  PWL(...)      gap current from the field solver
        B1  src 0    V = vs(t)     source voltage (none: 0)
        R   src out  R
        Cp  out 0    C             the port's grid capacitance

    Node equation: C dV/dt = I + (vs(t) - V) / R.  Over each step
    [(n-1) dt, n dt] the current keeps the value I_n handed over at step n
    (a one-point PWL is constant) and vs is taken linear between its end
    values; the linear equation is then integrated exactly in FP64:
    V(h) = V0 E + g0 A + k tau (h - A), E = exp(-h/tau), A = tau (1 - E),
    tau = R C, g0 = (I + vs0/R)/C, k = (vs1 - vs0)/(R C h).  Returns the
    node voltage V(n dt) of each port."""

    def __init__(self, dt: float, resistance: Sequence[float], capacitance: Sequence[float],
                 sources: Optional[Sequence[Optional[Callable[[float], float]]]] = None):
        self.dt = float(dt)
        self.R = np.asarray(resistance, np.float64)
        self.C = np.asarray(capacitance, np.float64)
        self.src = list(sources) if sources is not None else [None] * self.R.size
        self.V = np.zeros(self.R.size)

    def vs(self, q: int, t: float) -> float:
        f = self.src[q]
        return 0.0 if f is None else float(f(t))

    def step(self, n: int, currents) -> np.ndarray:
        h = self.dt
        t0, t1 = (n - 1) * h, n * h
        out = np.empty(self.R.size)
        for q in range(self.R.size):
            R, C, I = self.R[q], self.C[q], float(currents[q])
            tau = R * C
            E = math.exp(-h / tau)
            A = tau * (1.0 - E)
            v0, v1 = self.vs(q, t0), self.vs(q, t1)
            g0 = (I + v0 / R) / C
            k = (v1 - v0) / (R * C * h)
            self.V[q] = self.V[q] * E + g0 * A + k * tau * (h - A)
            out[q] = self.V[q]
        return out


def gaussian_sine(amplitude: float, f0: float, t0: float, width: float) -> Callable[[float], float]:
    """vs(t) = a sin(2 pi f0 t) exp(-ln2 ((t - t0)/width)^2), the behavioural
    source of the reference's SPICE example netlist."""
    return lambda t: amplitude * math.sin(2.0 * math.pi * f0 * t) * math.exp(-math.log(2.0) * ((t - t0) / width) ** 2)


def run_cosim(sim, circuit, steps: Optional[int] = None) -> None:
    """Advance ``sim`` (a Simulation, or anything with the same
    ``port_exchange``) by ``steps`` steps (default: all that remain), handing
    each step's SPICE gap currents to ``circuit.step(n, currents)`` and its
    node voltages back (ports, then extra nodes)."""
    done = sim.steps_done
    steps = sim.plan.time_steps - done if steps is None else int(steps)
    if steps <= 0:
        return
    cur = sim.port_exchange(None)
    for q in range(1, steps + 1):
        volts = circuit.step(done + q, cur)
        cur = sim.port_exchange(volts, current=q < steps)
