"""torch.autograd bridge for the far-field design gradient.

:func:`far_field_power` maps a tensor of per-cell relative permittivity of a
design box to J = |r E_pol(theta, phi, f)|^2 of a full simulation, and its
backward pass returns dJ/d eps from the discrete adjoint
(aem_adjoint_design_gradient).  The forward pass runs the deck on the HIP
device and keeps the simulation -- with its design-box record -- alive until
backward, so one forward + backward costs one recorded forward run and one
adjoint sweep.  PyTorch is plumbing here: the simulation and the gradient run
in libaem.

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
