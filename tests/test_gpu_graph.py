"""A step's device phase as one HIP graph launch (bh_step_graph).

The captured step must reproduce the eager step on every replay: the graph
is captured once per photon-store parity and then replayed, with the host's
own eager work (step physics, mesh and energy uploads, census export)
between the replays.
"""
import json
import os

import numpy as np
import pytest

from branson_amd import capi

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _close6(a, b):
    return abs(a - b) <= 5.01e-6 * abs(b) if b != 0.0 else abs(a) < 1e-300


def test_graph_driver_reproduces_reference_marshak_run():
    run = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["runs"]["marshak_slab_100"]
    with capi.Driver(os.path.join(GOLDEN, run["deck"]), algorithm=capi.ALGO_PERSISTENT,
                     step_graph=True) as d:
        reps = d.run()
    assert len(reps) == len(run["steps"])
    for r, g in zip(reps, run["steps"]):
        assert r["transported"] == g["transported"]
        assert r["census_size"] == g["census_size"]
        for k in ("absorbed_E", "exit_E", "post_census_E", "post_mat_E"):
            assert _close6(r[k], g[k]), (r["step"], k, r[k], g[k])


@pytest.mark.parametrize("source_mode", [capi.SOURCE_AUTO, capi.SOURCE_IN_TRANSPORT])
@pytest.mark.parametrize("prec", [capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK])
def test_graph_steps_equal_eager_steps(prec, source_mode):
    """source_mode SOURCE_IN_TRANSPORT: the graph's new photons are made
    inside the transport kernel, the eager steps' by the source kernel."""
    deck = os.path.join(GOLDEN, "decks", "cube_40.xml")
    out = {}
    for graph in (False, True):
        with capi.Driver(deck, algorithm=capi.ALGO_PERSISTENT, precision=prec, t_stop=0.005,
                         step_graph=graph, source_mode=source_mode if graph else 0) as d:
            steps = []
            while not d.finished():
                d.prepare()
                d.transport()
                r = d.finish()
                steps.append((r, d.tallies()[0].copy()))
                # eager work between graph replays
                e = capi.lib()
                n = capi.C.c_uint64()
                assert e.bh_census_size(e.bh_driver_engine(d._d), capi.C.byref(n)) == 0
            out[graph] = steps
    assert len(out[True]) == len(out[False]) == 5
    for (re, ae), (rg, ag) in zip(out[False], out[True]):
        assert rg["transported"] == re["transported"]
        assert rg["census_size"] == re["census_size"]
        assert rg["n_segments"] == re["n_segments"]
        assert np.allclose(ag, ae, rtol=1e-12, atol=0)
        assert abs(rg["post_census_E"] - re["post_census_E"]) <= 1e-12 * re["post_census_E"]


def _census_export(d):
    """Eager kernel work on the driver's engine: the census in AOS form."""
    L = capi.lib()
    h = L.bh_driver_engine(d._d)
    n = capi.C.c_uint64()
    assert L.bh_census_size(h, capi.C.byref(n)) == 0
    out = np.zeros(n.value, dtype=capi.PHOTON_DTYPE)
    got = capi.C.c_uint64()
    assert L.bh_census_export(h, out.ctypes.data, n.value, capi.C.byref(got)) == 0
    return out


def test_graph_replays_with_census_exports_between_them():
    deck = os.path.join(GOLDEN, "decks", "cube_40.xml")
    cens = {}
    for graph in (False, True):
        with capi.Driver(deck, algorithm=capi.ALGO_PERSISTENT, t_stop=0.004, step_graph=graph,
                         n_photons=200_000) as d:
            cens[graph] = []
            while not d.finished():
                r = d.step()
                scale = max(r["post_mat_E"], r["absorbed_E"] + r["exit_E"] + r["post_census_E"])
                assert abs(r["rad_conservation"]) < 1e-12 * scale
                cens[graph].append(_census_export(d))
    # Events are exact.  Positions carry round-off from the step physics: the
    # tallies are sums of device atomics in a run-dependent order, so the next
    # step's temperatures and opacities differ in the last bits between any
    # two runs, eager or graph, and flight lengths with them.  Two eager runs
    # of this deck differ by up to 1e-14 cm after four steps (cube of 2 cm).
    for ce, cg in zip(cens[False], cens[True]):
        assert len(cg) == len(ce) > 0
        for f in ("cell", "group", "rng", "descriptor"):
            assert np.array_equal(cg[f], ce[f]), f
        assert np.abs(cg["pos"] - ce["pos"]).max() <= 1e-12
