"""Array-built problems (branson_amd.problems): mesh connectivity, the
correlated lognormal field, and the step physics, which must equal the
oracle's Mesh::calculate_photon_energy restatement bit for bit."""
import os

import numpy as np

import oracle
from branson_amd import capi, problems

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def test_box_mesh_connectivity():
    nx, ny, nz = 5, 4, 3
    nodes, bc, nxt = problems.box_mesh(nx, ny, nz, 0.1, 0.2, 0.3, capi.VACUUM)
    n = nx * ny * nz
    assert nodes.shape == (n, 6) and bc.shape == (n, 6) and nxt.shape == (n, 6)
    for c in range(n):
        for face in range(6):
            if bc[c, face] == capi.ELEMENT:
                o = nxt[c, face]
                # the neighbour points back through the opposite face and
                # shares the node coordinate exactly
                assert nxt[o, face ^ 1] == c
                assert nodes[c, face] == nodes[o, face ^ 1]
            else:
                assert bc[c, face] == capi.VACUUM
    assert (bc == capi.ELEMENT).sum() == 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz
                                              + nx * ny * (nz - 1))
    # x fastest
    assert nodes[1, 0] == 0.1 and nodes[nx, 2] == 0.2 and nodes[nx * ny, 4] == 0.3


def test_lognormal_field_statistics():
    f = problems.lognormal_field((32, 32, 32), 10.0, 0.7, 4.0, 11)
    lz = np.log(f / 10.0) / 0.7
    # geometric mean is the median and the log-field has unit variance, exactly
    assert abs(np.exp(np.log(f).mean()) - 10.0) < 1e-9
    assert abs(lz.std() - 1.0) < 1e-12
    # Gaussian filter of std s: correlation exp(-lag^2 / (4 s^2))
    z = lz.reshape(32, 32, 32)
    for lag in (1, 4):
        r = np.mean(z * np.roll(z, lag, axis=2))
        assert abs(r - np.exp(-lag * lag / 64.0)) < 0.08
    assert not np.array_equal(f, problems.lognormal_field((32, 32, 32), 10.0, 0.7, 4.0, 12))


def test_step_physics_equals_oracle():
    sim = oracle.Sim(oracle.Deck.from_xml(os.path.join(DECKS, "cube_40.xml")))
    sim.prepare()
    cells, ph = sim.cells(), sim.physics()
    p = problems.CellProblem(cells.nodes, cells.bc, cells.next, 5.0, 1.0, 1.0, 0.001,
                             cv=0.1, rho=1.0)
    assert np.array_equal(p.fleck, cells.fleck)
    assert np.array_equal(p.emission_E, ph["emission_E"])


def test_with_params_keeps_the_source():
    p = problems.adjoint_field_problem(n=8)
    q = p.with_params(sigma_a=p.sigma_a * 2.0)
    assert np.array_equal(q.emission_E, p.emission_E)
    assert np.array_equal(q.sigma_a, 2.0 * p.sigma_a)
    assert np.array_equal(q.fleck, p.fleck)


def test_batch_concatenation_keeps_instances_apart():
    b = problems.batched_instances(n_instances=10, seed=1, cells_range=(20, 400),
                                   photons_range=(100, 1000))
    inst = b.instance_of_cell()
    assert len(inst) == b.n_cells == b.first_cell[-1]
    inner = b.bc == capi.ELEMENT
    rows, faces = np.nonzero(inner)
    # every neighbour lies in the same instance
    assert np.array_equal(inst[b.next[rows, faces]], inst[rows])
    assert np.array_equal(b.seeds, np.arange(10))
    sizes = np.diff(b.first_cell)
    assert np.all(sizes > 0)
