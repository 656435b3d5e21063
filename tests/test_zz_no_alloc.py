"""bh_transport_photons_noalloc against bh_transport_photons, through the C-ABI.

The entry point that allocates no device memory runs the same transport
kernels on buffers carved from the library's static scratch array.  On the
same photons it must give the existing entry point's results: every event,
cell, group and RNG counter bit for bit, positions and weights bit for bit
(the same arithmetic per photon), and per-cell tallies within 1e-12 relative
(its tally form adds into the caller's arrays on the device, so the
summation order differs).  Two calls in a row, the second with fewer
photons, check that the scratch left by the first call does not leak into
the second.
"""
import os

import numpy as np
import pytest

import oracle
from branson_amd import capi

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _photons(deck, n_user):
    d = oracle.Deck.from_xml(os.path.join(DECKS, deck))
    sim = oracle.Sim(d)
    sim.prepare()
    cells, ph, clk = sim.cells(), sim.physics(), sim.clock()
    photons = oracle.make_photons(cells, ph["emission_E"], ph["source_E"], ph["source_face"],
                                  n_user, ph["total_E"], d.seed, 1, 0, clk["dt"])
    return cells, photons


def _same(got, ref, a_g, t_g, a_r, t_r, st_g, st_r):
    for f in ("cell", "group", "source_type", "rng", "pos", "angle", "E", "E0", "life_dx"):
        assert np.array_equal(got[f], ref[f]), f
    assert np.array_equal(got["descriptor"], ref["descriptor"])
    for k in ("n_transported", "n_segments", "n_flushes", "n_census", "n_exit", "n_killed",
              "n_replayed"):
        assert st_g[k] == st_r[k], k
    assert st_g["census_E"] == st_r["census_E"] and st_g["exit_E"] == st_r["exit_E"]
    assert np.allclose(a_g, a_r, rtol=1e-12, atol=1e-300)
    assert np.allclose(t_g, t_r, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("kw", [dict(algorithm=capi.ALGO_PERSISTENT),
                                dict(algorithm=capi.ALGO_HISTORY),
                                dict(algorithm=capi.ALGO_PERSISTENT,
                                     precision=capi.PRECISION_FP32_TRACK)],
                         ids=["fp64-persistent", "fp64-history", "fp32-persistent"])
@pytest.mark.parametrize("deck,n_user", [("cube_40.xml", 120_000),
                                         ("marshak_slab_100.xml", 60_000)])
def test_noalloc_equals_transport_photons_twice(deck, n_user, kw):
    cells, photons = _photons(deck, n_user)
    n1 = len(photons)
    n2 = n1 // 3  # the second call is smaller
    print(f"noalloc: {deck} n1={n1} n2={n2} {kw}", flush=True)
    with capi.Engine(0, **kw) as e:
        e.set_mesh_cells(cells)
        for sl in (slice(0, n1), slice(n1 - n2, n1)):
            ref = np.ascontiguousarray(photons[sl])
            got = ref.copy()
            a_r, t_r, st_r = e.transport_photons(ref)
            a_g, t_g, st_g = e.transport_photons_noalloc(got)
            _same(got, ref, a_g, t_g, a_r, t_r, st_g, st_r)
            assert st_g["n_transported"] == len(ref)


def test_noalloc_rejects_a_call_beyond_its_scratch():
    cells, photons = _photons("cube_40.xml", 20_000)
    big = np.resize(photons, 600_000)  # 600k photons need ~140 MiB: beyond 64 MiB
    with capi.Engine(0) as e:
        e.set_mesh_cells(cells)
        with pytest.raises(capi.BhError) as ei:
            e.transport_photons_noalloc(big)
        assert "scratch" in str(ei.value)
        # the engine still works afterwards, and so does the entry point
        ref = photons.copy()
        got = photons.copy()
        a_r, t_r, st_r = e.transport_photons(ref)
        a_g, t_g, st_g = e.transport_photons_noalloc(got)
        _same(got, ref, a_g, t_g, a_r, t_r, st_g, st_r)


def test_noalloc_equals_transport_photons_on_a_batch():
    """A batch of twelve problems in one engine, each a lattice (per-instance
    lattice records, each photon's seed from its instance): the same
    agreement, twice, the second call smaller."""
    from branson_amd import problems
    b = problems.batched_instances(n_instances=12, seed=4, cells_range=(20, 800),
                                   photons_range=(1000, 20000))
    em = problems.emission_energy(b.nodes, b.cat("sigma_a"), b.cat("fleck"), b.cat("T"),
                                  b.parts[0].dt)
    total_E = np.bincount(b.instance_of_cell(), weights=em, minlength=b.n_instances)
    n = b.n_cells
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        e.set_mesh(b.nodes, b.bc, b.next, b.cat("sigma_a"), b.cat("sigma_s"), b.cat("fleck"))
        e.set_instances(b.first_cell, b.seeds)
        e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8), b.n_photons,
                           total_E, 1, 0, b.parts[0].dt)
        photons = e.photons()
        n1 = len(photons)
        n2 = n1 // 3
        print(f"noalloc batch: n1={n1} n2={n2}", flush=True)
        for sl in (slice(0, n1), slice(n1 - n2, n1)):
            ref = np.ascontiguousarray(photons[sl])
            got = ref.copy()
            a_r, t_r, st_r = e.transport_photons(ref)
            a_g, t_g, st_g = e.transport_photons_noalloc(got)
            assert st_r["lattice"] == 2 and st_g["lattice"] == 2
            _same(got, ref, a_g, t_g, a_r, t_r, st_g, st_r)
