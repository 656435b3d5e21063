"""Replicated time steps of many independent problems in one engine.

Every instance of a problems.Batch keeps its own material state; each step
runs Branson's step physics for all cells at once on the host
(Mesh::calculate_photon_energy and Mesh::update_temperature,
src/mesh.h:265-281 and :327-360, emission-only problems), then ONE batched
source (bh_source_instances) and ONE transport launch over the photons of all
instances, which writes their census on the device.  Instances of very
different sizes therefore share the GPU: the persistent transport kernel
refills lanes from whichever instance still has photons.
"""
from __future__ import annotations

import time

import numpy as np

from . import capi, problems


class BatchedRun:
    def __init__(self, batch: "problems.Batch", engine: "capi.Engine"):
        self.b = batch
        self.e = engine
        self.inst = batch.instance_of_cell()
        self.sigma_a = batch.cat("sigma_a")
        self.sigma_s = batch.cat("sigma_s")
        self.T = batch.cat("T")
        self.cv = np.concatenate([np.full(p.n_cells, p.cv) for p in batch.parts])
        self.rho = np.concatenate([np.full(p.n_cells, p.rho) for p in batch.parts])
        self.dt = batch.parts[0].dt
        nd = batch.nodes
        self.vol = (nd[:, 1] - nd[:, 0]) * (nd[:, 3] - nd[:, 2]) * (nd[:, 5] - nd[:, 4])
        self.step_no = 1
        self.census_E = 0.0
        self._prepare()
        engine.set_instances(batch.first_cell, batch.seeds)

    def _prepare(self):
        """Fleck factors and emission energies of the coming step."""
        b = self.b
        self.fleck = problems.fleck_factor(self.sigma_a, self.T, self.dt, self.cv, self.rho)
        self.emission_E = problems.emission_energy(b.nodes, self.sigma_a, self.fleck, self.T,
                                                   self.dt)
        self.total_E = np.bincount(self.inst, weights=self.emission_E,
                                   minlength=b.n_instances)
        self.e.set_mesh(b.nodes, b.bc, b.next, self.sigma_a, self.sigma_s, self.fleck)

    def step(self) -> dict:
        """One step of every instance; returns the step's counters, the
        device-phase seconds and the energy balance."""
        b, e = self.b, self.e
        n = b.n_cells
        t0 = time.perf_counter()
        e.source_instances(self.emission_E, np.zeros(n), None, np.full(n, -1, dtype=np.int8),
                           b.n_photons, self.total_E, self.step_no, 0, self.dt)
        abs_E, track_E, st = e.transport(self.dt)
        t1 = time.perf_counter()
        # Mesh::update_temperature, emission-only problems
        self.T = self.T + (abs_E - self.emission_E) / (self.cv * self.vol * self.rho)
        balance = (abs_E.sum() + st["census_E"] + st["exit_E"]) - \
                  (self.emission_E.sum() + self.census_E)
        self.census_E = st["census_E"]
        rep = dict(st, step=self.step_no, device_s=t1 - t0, balance=balance,
                   abs_E=abs_E, emission_E_total=float(self.emission_E.sum()))
        self.step_no += 1
        self._prepare()
        return rep
