"""Differentiable absorbed-energy tally: torch.autograd over the C-ABI.

    problem = FixedSource.sample(engine, cell_problem, n_photons=10**6, seed=7)
    abs_E = absorbed_energy(sigma_a, fleck, problem)      # torch float64
    loss = ((abs_E - target) ** 2).sum()
    loss.backward()                                        # sigma_a.grad, fleck.grad

The forward pass transports a copy of the problem's source photons
(bh_transport_photons) with the given per-cell opacities and Fleck factors;
the backward pass is the path-replay adjoint (bh_adjoint_photons): the
derivative of the expected tally with the source photons held fixed, each
history replayed from its Random123 counter.  Both run on the HIP engine;
there is no CPU path.
"""
from __future__ import annotations

import numpy as np

from . import capi


class FixedSource:
    """A mesh (geometry, scattering opacities) and a fixed set of source
    photons in their pre-transport state, on one engine."""

    def __init__(self, engine: "capi.Engine", nodes, bc, next_, sigma_s, photons: np.ndarray):
        self.engine = engine
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.bc = np.ascontiguousarray(bc, dtype=np.int32)
        self.next = np.ascontiguousarray(next_, dtype=np.uint32)
        self.n_cells = self.nodes.shape[0]
        self.sigma_s = np.ascontiguousarray(sigma_s, dtype=np.float64).reshape(self.n_cells, -1)
        assert photons.dtype == capi.PHOTON_DTYPE
        self.photons = np.ascontiguousarray(photons)

    @classmethod
    def sample(cls, engine, problem, n_photons: int, seed: int, cycle: int = 1):
        """Emission photons of a problems.CellProblem, sampled on the device
        exactly as a replicated step's source (bh_source)."""
        engine.set_mesh(problem.nodes, problem.bc, problem.next, problem.sigma_a,
                        problem.sigma_s, problem.fleck)
        n = problem.n_cells
        engine.source(problem.emission_E, np.zeros(n), None, np.full(n, -1, dtype=np.int8),
                      n_photons, float(problem.emission_E.sum()), seed, cycle, 0, problem.dt)
        return cls(engine, problem.nodes, problem.bc, problem.next, problem.sigma_s,
                   engine.photons())

    def _set(self, sigma_a, fleck):
        sa = np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(self.n_cells, 1)
        G = self.sigma_s.shape[1]
        self.engine.set_mesh(self.nodes, self.bc, self.next, np.repeat(sa, G, axis=1),
                             self.sigma_s, np.ascontiguousarray(fleck, dtype=np.float64))

    def absorbed(self, sigma_a, fleck, photons: np.ndarray | None = None) -> np.ndarray:
        """abs_E per cell (numpy) for these parameters."""
        self._set(sigma_a, fleck)
        ph = (self.photons if photons is None else photons).copy()
        a, _, _ = self.engine.transport_photons(ph)
        return a

    def gradient(self, sigma_a, fleck, abs_adjoint, photons: np.ndarray | None = None):
        """(dL/dsigma_a, dL/dfleck) per cell for abs_adjoint = dL/d abs_E."""
        self._set(sigma_a, fleck)
        ph = self.photons if photons is None else photons
        da, df, _ = self.engine.adjoint_photons(ph, abs_adjoint)
        return da, df


def _torch():
    import torch
    return torch


class _AbsorbedEnergy:
    """Holder for the lazily created torch.autograd.Function subclass."""
    fn = None


def _function():
    if _AbsorbedEnergy.fn is not None:
        return _AbsorbedEnergy.fn
    torch = _torch()

    class AbsorbedEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, sigma_a, fleck, problem):
            sa = sigma_a.detach().to("cpu", torch.float64).numpy()
            f = fleck.detach().to("cpu", torch.float64).numpy()
            ctx.problem = problem
            ctx.save_for_backward(sigma_a, fleck)
            out = problem.absorbed(sa, f)
            return torch.from_numpy(out).to(sigma_a.device)

        @staticmethod
        def backward(ctx, grad_abs):
            sigma_a, fleck = ctx.saved_tensors
            g = grad_abs.detach().to("cpu", torch.float64).numpy()
            da, df = ctx.problem.gradient(sigma_a.detach().to("cpu", torch.float64).numpy(),
                                          fleck.detach().to("cpu", torch.float64).numpy(), g)
            return (torch.from_numpy(da).to(sigma_a.device, sigma_a.dtype),
                    torch.from_numpy(df).to(fleck.device, fleck.dtype), None)

    _AbsorbedEnergy.fn = AbsorbedEnergy
    return AbsorbedEnergy


def absorbed_energy(sigma_a, fleck, problem: FixedSource):
    """Differentiable abs_E(sigma_a, fleck) for the problem's fixed source
    photons (torch tensors of n_cells entries in, n_cells out)."""
    return _function().apply(sigma_a, fleck, problem)
