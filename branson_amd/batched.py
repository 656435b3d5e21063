"""Replicated time steps of many independent problems in one engine.

Every instance of a problems.Batch keeps its own material state on the
device.  A step is one call, bh_step_resident_instances: the step physics of
every cell (Mesh::calculate_photon_energy and Mesh::update_temperature,
src/mesh.h:237-423; here emission-only problems whose absorption opacity is a
per-cell constant), each instance's photon energy, ONE batched source, ONE
transport launch over the photons of all instances, which writes their
census, and the per-instance energy balance.  Instances of very different
sizes therefore share the GPU: the persistent transport kernel refills lanes
from whichever instance still has photons.  Steps after the first are one
HIP graph launch each.
"""
from __future__ import annotations

import time

import numpy as np

from . import capi, problems


class BatchedRun:
    def __init__(self, batch: "problems.Batch", engine: "capi.Engine"):
        self.b = batch
        self.e = engine
        b = batch
        n = b.n_cells
        self.dt = b.parts[0].dt
        sigma_a = b.cat("sigma_a")
        sigma_s = b.cat("sigma_s")
        cv = np.concatenate([np.full(p.n_cells, p.cv) for p in b.parts])
        rho = np.concatenate([np.full(p.n_cells, p.rho) for p in b.parts])
        engine.set_mesh(b.nodes, b.bc, b.next, sigma_a, sigma_s, b.cat("fleck"))
        engine.set_instances(b.first_cell, b.seeds)
        # opacity law op_a = A + B T^C with B = 0: the per-cell constant
        # sigma_a; no initial radiation, no source faces
        engine.set_material(sigma_a, 0.0, 0.0, sigma_s, cv, rho, b.cat("T"), 0.0, 0.0,
                            np.full(n, -1, dtype=np.int8))
        self.step_no = 1

    def step(self) -> dict:
        """One step of every instance; returns the step's stats, the
        device-phase seconds, the per-instance reports and the energy
        balance (overall, and the largest per-instance relative one)."""
        t0 = time.perf_counter()
        tot, per = self.e.step_resident_instances(self.b.n_photons, self.step_no, 0, self.dt,
                                                  self.dt)
        t1 = time.perf_counter()
        st = tot["stats"]
        lhs = per["absorbed_E"] + per["post_census_E"] + per["exit_E"]
        rhs = per["pre_census_E"] + per["emission_E"] + per["source_E"]
        scale = np.maximum(np.maximum(lhs, rhs), 1e-300)
        rep = dict(st, step=self.step_no, device_s=t1 - t0, graph=tot["graph"],
                   balance=float(lhs.sum() - rhs.sum()),
                   max_instance_balance=float(np.max(np.abs(lhs - rhs) / scale)),
                   instances=per, emission_E_total=float(per["emission_E"].sum()))
        self.step_no += 1
        return rep

    def tallies(self):
        return self.e.tallies()
