"""Problems built from arrays rather than a Branson XML deck.

Branson's deck gives each cell a region through division triples (at most
1000 divisions per axis, src/input.h), so per-cell material fields and many
independent problems are built here directly as the cell arrays of the C-ABI
(bh_cells): nodes, boundary kinds, neighbours, opacities and Fleck factors.
The step physics follows Mesh::calculate_photon_energy (src/mesh.h:265-281)
operation for operation.
"""
from __future__ import annotations

import numpy as np

from .capi import ELEMENT, REFLECT, VACUUM

A_RAD = 0.01372      # GJ / cm^3 / keV^4 (src/constants.h:22)
C_LIGHT = 299.792458  # cm / sh (src/constants.h:17)


def box_mesh(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float,
             bc: int = VACUUM, origin=(0.0, 0.0, 0.0)):
    """A rectilinear nx*ny*nz box of uniform cells, x fastest.  Returns
    (nodes[n,6], bc[n,6], next[n,6]); outer faces get `bc` (or, if `bc` is a
    6-tuple, one kind per face X_NEG..Z_POS)."""
    kinds = tuple(bc) if np.ndim(bc) else (bc,) * 6
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest
    n = nx * ny * nz
    x0, y0, z0 = origin
    nodes = np.stack([x0 + i * dx, x0 + (i + 1) * dx, y0 + j * dy, y0 + (j + 1) * dy,
                      z0 + k * dz, z0 + (k + 1) * dz], axis=1)
    idx = np.arange(n)
    nxt = np.empty((n, 6), dtype=np.uint32)
    bcs = np.empty((n, 6), dtype=np.int32)
    for face, (axis, step, lim) in enumerate([(i, -1, 0), (i, 1, nx - 1), (j, -nx, 0),
                                               (j, nx, ny - 1), (k, -nx * ny, 0),
                                               (k, nx * ny, nz - 1)]):
        inner = axis != lim
        nxt[:, face] = np.where(inner, idx + step, idx)
        bcs[:, face] = np.where(inner, ELEMENT, kinds[face])
    return nodes, bcs, nxt


def lognormal_field(shape, median: float, sigma_log: float, corr_cells: float,
                    seed: int) -> np.ndarray:
    """A correlated lognormal field on a periodic grid: white noise filtered
    by a Gaussian of standard deviation corr_cells (in cells, by FFT),
    rescaled to unit variance, then median * exp(sigma_log * z).  Returned
    flattened x fastest (shape is (nx, ny, nz))."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(tuple(shape[::-1]))  # (nz, ny, nx): x fastest in memory
    kernel = np.ones_like(z)
    for ax, m in enumerate(z.shape):
        freq = np.fft.fftfreq(m)
        g = np.exp(-2.0 * (np.pi * freq * corr_cells) ** 2)
        shp = [1, 1, 1]
        shp[ax] = m
        kernel = kernel * g.reshape(shp)
    zf = np.real(np.fft.ifftn(np.fft.fftn(z) * kernel))
    zf = (zf - zf.mean()) / zf.std()
    return (median * np.exp(sigma_log * zf)).ravel()


def fleck_factor(sigma_a, T, dt, cv, rho):
    """f = 1 / (1 + dt op_a c (4 a T^3 / (cV rho))) (src/mesh.h:267)."""
    return 1.0 / (1.0 + dt * sigma_a * C_LIGHT * (4.0 * A_RAD * np.power(T, 3) / (cv * rho)))


def emission_energy(nodes, sigma_a, fleck, T, dt, replicated_factor=1.0):
    """replicated_factor dt vol f op_a a c T^4 (src/mesh.h:272-273)."""
    vol = (nodes[:, 1] - nodes[:, 0]) * (nodes[:, 3] - nodes[:, 2]) * (nodes[:, 5] - nodes[:, 4])
    return replicated_factor * dt * vol * fleck * sigma_a * A_RAD * C_LIGHT * np.power(T, 4)


class CellProblem:
    """One problem's cell arrays plus its step physics (emission only)."""

    def __init__(self, nodes, bc, next_, sigma_a, sigma_s, T, dt, cv=0.1, rho=1.0,
                 fleck=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.bc = np.ascontiguousarray(bc, dtype=np.int32)
        self.next = np.ascontiguousarray(next_, dtype=np.uint32)
        n = self.nodes.shape[0]
        self.n_cells = n
        self.sigma_a = np.broadcast_to(np.asarray(sigma_a, dtype=np.float64), (n,)).copy()
        self.sigma_s = np.broadcast_to(np.asarray(sigma_s, dtype=np.float64), (n,)).copy()
        self.T = np.broadcast_to(np.asarray(T, dtype=np.float64), (n,)).copy()
        self.dt, self.cv, self.rho = dt, cv, rho
        self.fleck = (fleck_factor(self.sigma_a, self.T, dt, cv, rho) if fleck is None
                      else np.broadcast_to(np.asarray(fleck, dtype=np.float64), (n,)).copy())
        self.emission_E = emission_energy(self.nodes, self.sigma_a, self.fleck, self.T, dt)

    def with_params(self, sigma_a=None, fleck=None):
        """The same geometry and source with other opacities / Fleck factors
        (the emission energies are kept: a fixed-source perturbation)."""
        p = object.__new__(CellProblem)
        p.__dict__.update(self.__dict__)
        if sigma_a is not None:
            p.sigma_a = np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(self.n_cells)
        if fleck is not None:
            p.fleck = np.ascontiguousarray(fleck, dtype=np.float64).reshape(self.n_cells)
        return p


def adjoint_field_problem(n: int = 32, width: float = 0.05, median: float = 10.0,
                          sigma_log: float = 0.7, corr_cells: float = 4.0,
                          sigma_s: float = 0.5, T: float = 0.5, dt: float = 0.001,
                          field_seed: int = 20240531) -> CellProblem:
    """The gradient workload: n^3 cells of `width` cm, vacuum faces, sigma_a a
    lognormal field (median 10/cm, sigma_log 0.7, correlation length 4
    cells), sigma_s = 0.5/cm, T = 0.5 keV, emission only (dt = 0.001 sh,
    cV = 0.1, rho = 1)."""
    nodes, bc, nxt = box_mesh(n, n, n, width, width, width, VACUUM)
    sa = lognormal_field((n, n, n), median, sigma_log, corr_cells, field_seed)
    return CellProblem(nodes, bc, nxt, sa, sigma_s, T, dt)


class Batch:
    """Independent problems concatenated into one mesh (bh_set_instances):
    instance k owns cells [first_cell[k], first_cell[k+1]), neighbours are
    global indices, and each instance keeps its own seed, photon count,
    temperature and opacities."""

    def __init__(self, parts, n_photons, seeds):
        self.parts = list(parts)
        self.n_instances = len(self.parts)
        self.n_photons = np.asarray(n_photons, dtype=np.uint64)
        self.seeds = np.asarray(seeds, dtype=np.uint32)
        sizes = [p.n_cells for p in self.parts]
        self.first_cell = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
        off = self.first_cell[:-1]
        self.nodes = np.concatenate([p.nodes for p in self.parts])
        self.bc = np.concatenate([p.bc for p in self.parts])
        self.next = np.concatenate([p.next.astype(np.int64) + o
                                    for p, o in zip(self.parts, off)]).astype(np.uint32)
        self.n_cells = int(self.first_cell[-1])

    def cat(self, attr):
        return np.concatenate([getattr(p, attr) for p in self.parts])

    def instance_of_cell(self):
        return np.repeat(np.arange(self.n_instances), np.diff(self.first_cell))


def batched_instances(n_instances: int = 256, seed: int = 0, cells_range=(50, 32768),
                      photons_range=(10**4, 10**6), width: float = 0.02,
                      dt: float = 0.01) -> Batch:
    """The batched workload: instances alternate between slabs (n x 1 x 1,
    reflecting side faces) and cubes (about n^(1/3) cells a side, vacuum),
    with cell counts and photon counts drawn log-uniform, per-cell sigma_a
    lognormal (median 10/cm, sigma_log 1.0), sigma_s uniform in [0, 2]/cm, and
    the instance temperature T0 uniform in [0.01, 1] keV.  Instance k uses
    seed k for its photons."""
    rng = np.random.default_rng(seed)
    lo_c, hi_c = np.log(cells_range[0]), np.log(cells_range[1])
    lo_p, hi_p = np.log(photons_range[0]), np.log(photons_range[1])
    parts, n_ph = [], []
    for k in range(n_instances):
        nc = int(round(np.exp(rng.uniform(lo_c, hi_c))))
        if k % 2 == 0:
            kinds = (VACUUM, VACUUM, REFLECT, REFLECT, REFLECT, REFLECT)
            nodes, bc, nxt = box_mesh(nc, 1, 1, width, width, width, kinds)
        else:
            m = max(2, int(round(nc ** (1.0 / 3.0))))
            nodes, bc, nxt = box_mesh(m, m, m, width, width, width, VACUUM)
        n = len(nodes)
        sa = 10.0 * np.exp(rng.normal(0.0, 1.0, n))
        ss = rng.uniform(0.0, 2.0, n)
        T0 = rng.uniform(0.01, 1.0)
        parts.append(CellProblem(nodes, bc, nxt, sa, ss, T0, dt))
        n_ph.append(int(round(np.exp(rng.uniform(lo_p, hi_p)))))
    return Batch(parts, n_ph, np.arange(n_instances))
