"""The path-replay adjoint on the GPU.

  * against the oracle's forward-mode differential-operator tally (the same
    estimator, computed history by history in the other direction): equal to
    round-off for random directions, on a vacuum cube, a reflecting slab lit
    through a source face, a 4-group problem and a batch of twelve problems
    in one engine;
  * against central finite differences with common random numbers
    (h = 1e-4 relative) on the 32^3 lognormal-opacity workload with 1e6
    photons: within statistical error, batch means over 20 photon batches;
  * through torch.autograd: backward() gives the direct C-ABI gradient.
"""
import math

import numpy as np
import pytest

import oracle
from branson_amd import capi, problems
from branson_amd.adjoint import FixedSource, absorbed_energy

pytestmark = pytest.mark.gpu


def _cells_from(p):
    return oracle.Cells(p.nodes, p.bc, p.next, p.sigma_a.reshape(-1, 1),
                        p.sigma_s.reshape(-1, 1), p.fleck)


def _case(kind):
    rng = np.random.default_rng({"cube": 1, "slab": 2, "groups": 3}[kind])
    if kind == "cube":
        nodes, bc, nxt = problems.box_mesh(6, 6, 6, 0.05, 0.05, 0.05, capi.VACUUM)
        n = len(nodes)
        cells = oracle.Cells(nodes, bc, nxt, np.exp(rng.normal(np.log(5), 0.5, n)).reshape(-1, 1),
                             np.full((n, 1), 1.0), rng.uniform(0.3, 0.8, n))
        em, src = np.ones(n), np.zeros(n)
        face = np.full(n, -1, dtype=np.int8)
    elif kind == "slab":
        kinds = (capi.SOURCE, capi.VACUUM, capi.REFLECT, capi.REFLECT, capi.REFLECT,
                 capi.REFLECT)
        nodes, bc, nxt = problems.box_mesh(20, 1, 1, 0.01, 0.01, 0.01, kinds)
        n = len(nodes)
        cells = oracle.Cells(nodes, bc, nxt, rng.uniform(20, 60, n).reshape(-1, 1),
                             np.full((n, 1), 10.0), rng.uniform(0.2, 0.9, n))
        em, src = np.full(n, 0.1), np.zeros(n)
        src[0] = 1.0
        face = np.full(n, -1, dtype=np.int8)
        face[0] = 0
    else:
        nodes, bc, nxt = problems.box_mesh(4, 4, 4, 0.05, 0.05, 0.05, capi.VACUUM)
        n = len(nodes)
        sa = np.exp(rng.normal(np.log(5), 0.5, n))
        cells = oracle.Cells(nodes, bc, nxt, np.repeat(sa[:, None], 4, axis=1),
                             rng.uniform(0.5, 3.0, (n, 4)), rng.uniform(0.3, 0.8, n))
        em, src = np.ones(n), np.zeros(n)
        face = np.full(n, -1, dtype=np.int8)
    photons = oracle.make_photons(cells, em, src, face, 60000, em.sum() + src.sum(), 99, 1, 0,
                                  0.001)
    return cells, photons, rng


@pytest.mark.parametrize("kind", ["cube", "slab", "groups"])
def test_adjoint_equals_oracle_differential_operator_tally(kind):
    cells, photons, rng = _case(kind)
    n = cells.n_cells
    g = rng.normal(size=n)
    with capi.Engine(0) as e:
        e.set_mesh_cells(cells)
        da, df, st = e.adjoint_photons(photons, g)
        # the adjoint leaves the caller's photons alone
        assert st["n_transported"] == len(photons)
    for _ in range(3):
        va, vf = rng.normal(size=n), rng.normal(size=n) * 0.1
        _, dabs, segs = oracle.dos_transport(cells, photons.copy(), va, vf, 8)
        ref = g @ dabs
        got = da @ va + df @ vf
        scale = np.abs(g * dabs).sum()
        print(kind, got, ref, scale, flush=True)
        assert abs(got - ref) <= 1e-9 * scale
    assert st["n_segments"] == segs


def test_adjoint_on_a_batch_equals_oracle_differential_operator_tally():
    """Twelve problems in one engine, each photon drawing from its own
    instance's seed: the adjoint equals the oracle's differential-operator
    tally on the concatenated mesh."""
    b = problems.batched_instances(n_instances=12, seed=4, cells_range=(20, 800),
                                   photons_range=(1000, 20000))
    n = b.n_cells
    em = problems.emission_energy(b.nodes, b.cat("sigma_a"), b.cat("fleck"), b.cat("T"),
                                  b.parts[0].dt)
    total_E = np.bincount(b.instance_of_cell(), weights=em, minlength=b.n_instances)
    cells = oracle.Cells(b.nodes, b.bc, b.next, b.cat("sigma_a").reshape(-1, 1),
                         b.cat("sigma_s").reshape(-1, 1), b.cat("fleck"))
    rng = np.random.default_rng(5)
    g = rng.normal(size=n)
    with capi.Engine(0) as e:
        e.set_mesh(b.nodes, b.bc, b.next, b.cat("sigma_a"), b.cat("sigma_s"), b.cat("fleck"))
        e.set_instances(b.first_cell, b.seeds)
        e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8), b.n_photons,
                           total_E, 1, 0, b.parts[0].dt)
        photons = e.photons()
        da, df, st = e.adjoint_photons(photons, g)
    assert st["n_transported"] == len(photons)
    for _ in range(2):
        va, vf = rng.normal(size=n), rng.normal(size=n) * 0.1
        _, dabs, segs = oracle.dos_transport(cells, photons.copy(), va, vf, 8)
        ref = g @ dabs
        got = da @ va + df @ vf
        assert abs(got - ref) <= 1e-9 * np.abs(g * dabs).sum()
    assert st["n_segments"] == segs


def test_adjoint_rejects_group_dependent_sigma_a():
    cells, photons, rng = _case("groups")
    sa = cells.sigma_a.copy()
    sa[:, 1] *= 1.5
    bad = oracle.Cells(cells.nodes, cells.bc, cells.next, sa, cells.sigma_s, cells.fleck)
    with capi.Engine(0) as e:
        e.set_mesh_cells(bad)
        with pytest.raises(capi.BhError):
            e.adjoint_photons(photons, np.zeros(cells.n_cells))


def test_adjoint_matches_central_differences_on_the_field_problem():
    prob = problems.adjoint_field_problem()  # 32^3, lognormal sigma_a
    n = prob.n_cells
    sa, f = prob.sigma_a, prob.fleck
    z = problems.lognormal_field((32, 32, 32), 1.0, 1.0, 4.0, 77)  # exp(z), smooth
    target_sa = sa * 1.05 * np.power(z, 0.1)  # mostly more absorbing: g one-signed
    # one-signed smooth directions, so the derivative is resolved
    va = sa * problems.lognormal_field((32, 32, 32), 1.0, 0.5, 4.0, 5)
    vf = -0.2 * f * problems.lognormal_field((32, 32, 32), 1.0, 0.5, 4.0, 6)
    h = 1e-4
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        src = FixedSource.sample(e, prob, 10**6, seed=2718)
        target = src.absorbed(target_sa, f)
        abs0 = src.absorbed(sa, f)
        g = 2.0 * (abs0 - target)  # L = sum (abs - target)^2
        adj, fd = [], []
        for b in np.array_split(np.arange(len(src.photons)), 20):
            ph = src.photons[b]
            da, df = src.gradient(sa, f, g, photons=ph)
            adj.append(da @ va + df @ vf)
            ap = src.absorbed(sa + h * va, f + h * vf, photons=ph)
            am = src.absorbed(sa - h * va, f - h * vf, photons=ph)
            fd.append(g @ (ap - am) / (2 * h))
    adj, fd = np.array(adj), np.array(fd)
    d = adj - fd
    se = d.std(ddof=1) / math.sqrt(len(d))
    z_score = d.mean() / se
    print("adjoint", adj.sum(), "fd", fd.sum(), "z", z_score, flush=True)
    assert abs(z_score) < 2.0
    # resolved derivative
    assert abs(adj.mean()) > 5 * adj.std(ddof=1) / math.sqrt(len(adj))


def test_autograd_backward_is_the_adjoint():
    torch = pytest.importorskip("torch")
    prob = problems.adjoint_field_problem(n=12)
    with capi.Engine(0) as e:
        src = FixedSource.sample(e, prob, 100000, seed=3)
        target = torch.from_numpy(src.absorbed(prob.sigma_a * 1.1, prob.fleck))
        sa = torch.tensor(prob.sigma_a, dtype=torch.float64, requires_grad=True)
        f = torch.tensor(prob.fleck, dtype=torch.float64, requires_grad=True)
        out = absorbed_energy(sa, f, src)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        # the forward pass is the transport (tally atomics: summation order varies)
        assert np.allclose(out.detach().numpy(), src.absorbed(prob.sigma_a, prob.fleck),
                           rtol=1e-12, atol=0)
        g = 2.0 * (out.detach().numpy() - target.numpy())
        da, df = src.gradient(prob.sigma_a, prob.fleck, g)
    assert np.allclose(sa.grad.numpy(), da, rtol=1e-12, atol=0)
    assert np.allclose(f.grad.numpy(), df, rtol=1e-12, atol=0)
    assert np.any(da != 0) and np.any(df != 0)
