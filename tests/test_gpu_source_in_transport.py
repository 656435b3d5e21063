"""New photons made inside the persistent transport kernel
(bh_options.source_mode): the photon each lane takes is sampled where it is
transported -- its draws computed by the whole wave, then the source's own
sampling -- instead of being written to HBM by the source kernel and read
straight back.  Against the source kernel's photons: every census photon
bit for bit and in order, the census energy identical, photon counts equal
and per-cell tallies within 1e-12 (device atomics add in another order), in
both precisions (the FP64 replay of the FP32 form makes a new photon again
from its draws), on a resident census carried over two steps."""
import os

import numpy as np
import pytest

import oracle
from branson_amd import capi

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "decks")


def _deck(name, n_photons, n_groups=1):
    path = (os.path.join(ROOT, "branson_amd", "decks", name) if name.startswith("hohlraum")
            else os.path.join(GOLDEN, name))
    d = oracle.Deck.from_xml(path, n_groups=n_groups)
    d.n_photons = n_photons
    d.t_stop = max(d.t_stop, d.t_start + 2.5 * d.dt_start)  # at least two steps
    return d


def _inputs(d, n_steps=2):
    """The oracle's per-step inputs (cells, physics, clock), computed once so
    that every engine run sees the same arrays (the oracle's threaded
    tallies, and with them the next step's temperatures, vary in the last
    bit from run to run)."""
    sim = oracle.Sim(d)
    out = []
    for _ in range(n_steps):
        sim.prepare()
        out.append((sim.cells(), sim.physics(), sim.clock()))
        sim.step(8)
    return out


def _two_steps(d, inputs, **kw):
    out = []
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT, **kw) as e:
        for cells, ph, clk in inputs:
            census_E = ph["census_E"] if clk["step"] == 1 else None
            e.set_mesh_cells(cells)
            n_new = e.source(ph["emission_E"], ph["source_E"], census_E, ph["source_face"],
                             d.n_photons, ph["total_E"], d.seed, clk["step"], 0, clk["dt"])
            a, t, st = e.transport(clk["next_dt"])
            out.append((n_new, a, t, st, e.census()))
    return out


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK])
@pytest.mark.parametrize("deck,n,G", [("cube_40.xml", 300_000, 1),
                                      ("marshak_slab_100.xml", 100_000, 1),
                                      ("hohlraum_64.xml", 1_000_000, 30)],
                         ids=["cube", "marshak", "hohlraum30"])
def test_photons_made_in_transport_equal_source_kernel(deck, n, G, precision):
    d = _deck(deck, n, G)
    kw = dict(precision=precision, fp32_tol=1e-6)
    inputs = _inputs(d)
    made = _two_steps(d, inputs, source_mode=capi.SOURCE_IN_TRANSPORT, **kw)
    stored = _two_steps(d, inputs, source_mode=capi.SOURCE_KERNEL, **kw)
    for (n1, a1, t1, s1, c1), (n2, a2, t2, s2, c2) in zip(made, stored):
        print(deck, precision, n1, s1["n_census"], s1["n_replayed"], flush=True)
        assert n1 == n2
        for k in ("n_transported", "n_census", "n_exit", "n_killed", "n_segments", "n_flushes",
                  "n_replayed"):
            assert s1[k] == s2[k], k
        assert s1["census_E"] == s2["census_E"]
        assert c1.tobytes() == c2.tobytes()
        assert np.allclose(a1, a2, rtol=1e-12, atol=1e-300)
        assert np.allclose(t1, t2, rtol=1e-12, atol=1e-300)


def test_photons_export_between_source_and_transport():
    """bh_photons_export writes photons that were left to the transport
    kernel into the store first: the export equals the oracle's source."""
    d = _deck("cube_40.xml", 200_000)
    sim = oracle.Sim(d)
    sim.prepare()
    cells, ph, clk = sim.cells(), sim.physics(), sim.clock()
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT,
                     source_mode=capi.SOURCE_IN_TRANSPORT) as e:
        e.set_mesh_cells(cells)
        e.source(ph["emission_E"], ph["source_E"], None, ph["source_face"], d.n_photons,
                 ph["total_E"], d.seed, 2, 0, clk["dt"])
        got = e.photons()
        a, t, st = e.transport(clk["next_dt"])
    ref = oracle.make_photons(cells, ph["emission_E"], ph["source_E"], ph["source_face"],
                              d.n_photons, ph["total_E"], d.seed, 2, 0, clk["dt"])
    for f in ("cell", "group", "source_type", "rng", "pos", "E", "E0", "life_dx"):
        assert np.array_equal(got[f], ref[f]), f
    assert st["n_transported"] == len(ref)
