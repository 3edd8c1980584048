"""torch.autograd bridge for the design gradients (far field and port S11).

:func:`far_field_power` maps a tensor of per-cell relative permittivity of a
design box to J = |r E_pol(theta, phi, f)|^2 of a full simulation, and its
backward pass returns dJ/d eps from the discrete adjoint
(aem_adjoint_design_gradient).  The forward pass runs the deck on the HIP
device and keeps the simulation -- with its design-box record -- alive until
backward, so one forward + backward costs one recorded forward run and one
adjoint sweep.  :func:`port_s11` does the same for S11 of a lumped port,
whose backward pass goes through the transposed port lines
(aem_adjoint_series).  PyTorch is plumbing here: the simulation and the
gradient run in libaem.

The reference's own workflow (optimization_scripts/adjoint_optimization/
adjoint_optimization.py:76-134) gets the gradient from a second forward run
with an adjoint source and FFT overlaps of field slices; this is the exact
gradient of the discrete update instead.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from .solver import Plan, Simulation


@dataclass
class DesignProblem:
    """A deck, its design box [lo, hi] (1-based cells, inclusive) and the
    far-field objective (theta, phi in degrees, frequency in Hz, 'theta' or
    'phi' polarisation).  The time-domain far-field series are off by
    default: the objective needs only the frequency-domain transform."""
    plan: Plan
    lo: Sequence[int]
    hi: Sequence[int]
    theta: float
    phi: float
    freq: float
    pol: str = "theta"
    ntff: str = "off"

    @property
    def shape(self):
        """(nz, ny, nx) of the design box."""
        return tuple(int(self.hi[a]) - int(self.lo[a]) + 1 for a in (2, 1, 0))

    def simulate(self, eps: np.ndarray) -> Simulation:
        """Run the full deck with the design permittivity ``eps`` and the
        objective on; the returned simulation can produce the gradient."""
        sim = Simulation(self.plan, ntff=self.ntff)
        sim.set_objective(self.theta, self.phi, self.freq, self.pol)
        sim.design_box(self.lo, self.hi)
        sim.set_design_eps(eps)
        sim.run()
        return sim


class _FarFieldPower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps: torch.Tensor, problem: DesignProblem) -> torch.Tensor:
        if tuple(eps.shape) != problem.shape:
            raise ValueError(f"eps must have the design box shape {problem.shape}, got {tuple(eps.shape)}")
        sim = problem.simulate(eps.detach().to("cpu", torch.float32).numpy())
        ctx.sim = sim
        ctx.meta = (eps.dtype, eps.device)
        F = sim.objective()
        return torch.tensor(abs(F) ** 2, dtype=eps.dtype, device=eps.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        sim = ctx.sim
        g = sim.design_gradient()
        sim.close()
        ctx.sim = None
        dtype, device = ctx.meta
        grad = torch.from_numpy(g).to(device=device, dtype=dtype) * grad_out
        return grad, None


def far_field_power(eps: torch.Tensor, problem: DesignProblem) -> torch.Tensor:
    """J = |r E_pol|^2 of ``problem`` with design permittivity ``eps`` (shape
    ``problem.shape``); differentiable in ``eps``."""
    return _FarFieldPower.apply(eps, problem)


@dataclass
class PortProblem:
    """A deck with lumped ports, its design box [lo, hi] (1-based cells,
    inclusive) and the frequency (Hz) of a port objective; ``port`` is the
    0-based port whose voltage spectrum enters (the excited port for S11)."""
    plan: Plan
    lo: Sequence[int]
    hi: Sequence[int]
    freq: float
    port: int = 0
    ntff: str = "off"

    @property
    def shape(self):
        """(nz, ny, nx) of the design box."""
        return tuple(int(self.hi[a]) - int(self.lo[a]) + 1 for a in (2, 1, 0))

    def simulate(self, eps: np.ndarray) -> Simulation:
        """Run and finalize the full deck with the design permittivity
        ``eps``; the returned simulation can produce the gradient."""
        sim = Simulation(self.plan, ntff=self.ntff)
        sim.design_box(self.lo, self.hi)
        sim.set_design_eps(eps)
        sim.run()
        sim.finalize()
        return sim


class _PortVoltage(torch.autograd.Function):
    """V(f0) and Vinc(f0); only V carries a gradient."""

    @staticmethod
    def forward(ctx, eps: torch.Tensor, problem: PortProblem):
        if tuple(eps.shape) != problem.shape:
            raise ValueError(f"eps must have the design box shape {problem.shape}, got {tuple(eps.shape)}")
        sim = problem.simulate(eps.detach().to("cpu", torch.float32).numpy())
        ctx.sim = sim
        ctx.problem = problem
        ctx.meta = (eps.dtype, eps.device)
        V = complex(sim.dft("port_v", [problem.freq], problem.port)[0])
        vinc = complex(sim.dft("incident", [problem.freq], 0)[0])
        return (torch.tensor(V, dtype=torch.complex128, device=eps.device),
                torch.tensor(vinc, dtype=torch.complex128, device=eps.device))

    @staticmethod
    def backward(ctx, grad_v: torch.Tensor, grad_inc: torch.Tensor):
        sim, problem = ctx.sim, ctx.problem
        g = complex(grad_v.item())        # dJ/dRe V + i dJ/dIm V
        grads = sim.port_gradient(problem.port, problem.freq, g)
        sim.close()
        ctx.sim = None
        dtype, device = ctx.meta
        return torch.from_numpy(grads).to(device=device, dtype=dtype), None


def port_s11(eps: torch.Tensor, problem: PortProblem) -> torch.Tensor:
    """S11 = V(f0)/Vinc(f0) - 1 of ``problem``'s port (post_processor.py:255-262)
    with design permittivity ``eps``, a complex tensor differentiable in
    ``eps``: the reference's adjoint objective (adjoint_optimization.py:76-80)
    is then ``(goal - port_s11(eps, problem)).abs() ** 2``.  The incident
    spectrum Vinc does not depend on the design."""
    V, vinc = _PortVoltage.apply(eps, problem)
    return V / vinc.detach() - 1
