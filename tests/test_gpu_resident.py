"""Device-resident step physics (bh_step_resident, driver device_physics):
opacities, Fleck factors, source energies, E_tot, temperatures and the
conservation sums computed on the device, with nothing but the step report
crossing to the host.

  * against the reference executable's printed per-step diagnostics
    (tests/golden/reference_runs.json): photon counts and census sizes
    identical at every step, energies to the 6 printed digits, conservation
    to round-off;
  * against the host-physics driver: the first step's per-cell opacities,
    Fleck factors and source energies within 1e-13 (device pow against glibc
    pow differ in the last bit), identical photon counts at every step, and
    the free-running bar of 1e-9 for later steps' energies and the final
    temperatures.
"""
import json
import os

import numpy as np
import pytest

from branson_amd import capi

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _close6(a, b):
    return abs(a - b) <= 5.01e-6 * abs(b) if b != 0.0 else abs(a) < 1e-300


def _rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), 1e-300)
    return float(np.max(d / s)) if d.size else 0.0


@pytest.mark.parametrize("algo,source_mode", [(capi.ALGO_HISTORY, 0), (capi.ALGO_PERSISTENT, 0),
                                              (capi.ALGO_PERSISTENT, capi.SOURCE_IN_TRANSPORT)])
@pytest.mark.parametrize("name", ["marshak_slab_100", "cube_40"])
def test_resident_driver_reproduces_reference_run(name, algo, source_mode):
    """Persistent: steps after the first are one graph launch each; with
    SOURCE_IN_TRANSPORT their new photons are made inside the transport
    kernel."""
    run = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["runs"][name]
    with capi.Driver(os.path.join(GOLDEN, run["deck"]), algorithm=algo,
                     device_physics=True, source_mode=source_mode) as d:
        reps = d.run()
    assert len(reps) == len(run["steps"])
    for r, g in zip(reps, run["steps"]):
        print(name, r["step"], r["transported"], r["census_size"], r["absorbed_E"],
              r["rad_conservation"], flush=True)
        assert r["transported"] == g["transported"]
        assert r["census_size"] == g["census_size"]
        for k in ("emission_E", "source_E", "absorbed_E", "exit_E", "pre_census_E",
                  "post_census_E", "pre_mat_E", "post_mat_E"):
            assert _close6(r[k], g[k]), (name, r["step"], k, r[k], g[k])
        scale = max(r["post_mat_E"], r["absorbed_E"] + r["exit_E"] + r["post_census_E"])
        assert abs(r["rad_conservation"]) < 1e-12 * scale
        assert abs(r["mat_conservation"]) < 1e-12 * scale


@pytest.mark.parametrize("deck,kw", [
    ("marshak_slab_100", {}),
    ("cube_40", {"n_photons": 2_000_000, "t_stop": 0.003}),
    ("hohlraum_64", {"n_groups": 4, "n_photons": 4_000_000, "t_stop": 0.003}),
    # rank 1 of 2: E_tot summed serially, small-energy reassignment
    ("marshak_slab_100", {"n_ranks": 2, "rank": 1}),
])
def test_resident_physics_equals_host_physics(deck, kw):
    """Step 1 has identical inputs on both sides, so the physics kernel is
    held to 1e-13 per cell (the last bit of pow), and the step's sums to
    the summation error bounds.  Later steps start from
    temperatures that already carry the run's round-off (tally atomics; in
    the Marshak slab, hundreds of grazing reflections per photon grow ulps
    to ~1e-12 within a few steps), so they are held to exact photon counts
    and the 1e-9 bar of the free-running parity test on every step total."""
    path = (os.path.join(ROOT, "branson_amd", "decks", deck + ".xml") if deck == "hohlraum_64"
            else os.path.join(GOLDEN, "decks", deck + ".xml"))
    out = {}
    for dev in (False, True):
        with capi.Driver(path, algorithm=capi.ALGO_PERSISTENT, device_physics=dev, **kw) as d:
            if dev:
                reps = [d.step()]
                ph1 = d.physics()  # the step just transported
            else:
                d.prepare()
                ph1 = d.physics()  # the step just prepared
                d.transport()
                reps = [d.finish()]
            while not d.finished():
                reps.append(d.step())
            T_e, T_r = d.temperature()
            out[dev] = (reps, ph1, T_e, T_r)
    (rh, ph_h, Th, Trh), (rd, ph_d, Td, Trd) = out[False], out[True]
    for k in ("sigma_a", "sigma_s", "fleck", "emission_E", "source_E"):
        assert _rel(ph_d[k], ph_h[k]) <= 1e-13, k
    assert len(rh) == len(rd) >= 3
    for i, (a, b) in enumerate(zip(rd, rh)):
        assert a["transported"] == b["transported"]
        assert a["census_size"] == b["census_size"]
        # step 1: the same per-cell values (to the last bit of pow) summed two
        # ways -- the host's serial loop and the device's tree -- so the sums
        # differ by at most the two summation error bounds, 2 (n + 64) u
        bound = 2.0 * (len(Th) + 64) * 2.0 ** -53 + 1e-13
        keys = ("emission_E", "source_E", "pre_census_E", "pre_mat_E") if i == 0 else ()
        for k in keys:
            assert _rel(a[k], b[k]) <= bound or a[k] == b[k], (i, k, a[k], b[k])
        for k in ("emission_E", "source_E", "pre_census_E", "pre_mat_E", "absorbed_E",
                  "post_census_E", "post_mat_E", "exit_E"):
            assert _rel(a[k], b[k]) <= 1e-9 or a[k] == b[k], (i, k, a[k], b[k])
    # Per cell, after many steps: the two runs' later inputs differ at
    # round-off (the step sums are taken in different orders), and in the
    # slab's long chains of grazing reflections that now and then flips an
    # event at a near tie, which moves part of one photon's weight from one
    # cell to its neighbour -- about 1e-6 of a cell's temperature, while
    # every step's totals above stay within 1e-9.
    assert _rel(Td, Th) <= 1e-5
    assert _rel(Trd, Trh) <= 1e-5


def test_resident_engine_entry_points():
    """bh_set_material / bh_step_resident / bh_get_temperature directly: a
    box lit through one face, two steps; energy balance and a cooler source
    face than the wall it heats are checked, and bad input is rejected."""
    from branson_amd import problems
    n = 20
    kinds = (capi.SOURCE,) + (capi.VACUUM,) * 5
    nodes, bc, nxt = problems.box_mesh(n, 4, 4, 0.05, 0.05, 0.05, kinds)
    m = len(nodes)
    face = np.where(nodes[:, 0] == 0.0, 0, -1).astype(np.int8)
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        e.set_mesh(nodes, bc, nxt, np.ones((m, 1)), np.ones((m, 1)), np.full(m, 0.5))
        with pytest.raises(capi.BhError):
            e.step_resident(100_000, 7, 1, 0, 0.001, 0.001)  # no material yet
        with pytest.raises(capi.BhError):
            e.set_material(5.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.01, 0.0, 0.3, face)  # cv = 0
        e.set_material(5.0, 100.0, -3.0, 1.0, 0.1, 1.0, 0.01, 0.0, np.where(face == 0, 0.3, 0.0),
                       face)
        pre = None
        for step in (1, 2):
            r = e.step_resident(100_000, 7, step, 0, 0.001, 0.001)
            st = r["stats"]
            pre_census = r["pre_census_E"] if step == 1 else pre
            lhs = r["absorbed_E"] + st["census_E"] + st["exit_E"]
            rhs = pre_census + r["emission_E"] + r["source_E"]
            assert abs(lhs - rhs) <= 1e-12 * rhs
            mat = r["post_mat_E"] - (r["pre_mat_E"] + r["absorbed_E"] - r["emission_E"])
            assert abs(mat) <= 1e-12 * r["post_mat_E"]
            assert r["n_new"] > 0 and st["n_transported"] >= r["n_new"]
            assert r["etot_serial"] == 0  # one rank: parallel E_tot, counts proven
            pre = st["census_E"]
        T_e, T_r = e.temperature()
    assert np.all(np.isfinite(T_e)) and T_e[face == 0].mean() > T_e[face != 0].mean()


@pytest.mark.parametrize("deck,kw", [
    ("marshak_slab_100", dict(n_ranks=2, rank=1)),
    ("marshak_slab_100", dict(n_ranks=3, rank=2)),
    ("cube_40", dict(n_ranks=4, rank=3, n_photons=20_000)),
], ids=["marshak-1of2", "marshak-2of3", "cube-sparse-3of4"])
def test_resident_rank_share_matches_oracle(deck, kw):
    """A replicated rank's first step with its physics on the device
    (small-energy reassignment and all-rank E_tot, src/mesh.h:287-315,
    src/replicated_driver.h:57-59) against the oracle's restatement for the
    same rank: per-cell energies to the last bit of pow (1e-13), E_tot to
    its summation order, photon counts, census size and energies."""
    import oracle
    path = os.path.join(GOLDEN, "decks", deck + ".xml")
    with capi.Driver(path, algorithm=capi.ALGO_PERSISTENT, device_physics=True, **kw) as d:
        rep = d.step()
        ph = d.physics()
    od = oracle.Deck.from_xml(path)
    if "n_photons" in kw:
        od.n_photons = kw["n_photons"]
    sim = oracle.Sim(od)
    sim.set_rank(kw["rank"], kw["n_ranks"])
    sim.prepare()
    oph = sim.physics()
    orep = sim.step(8)
    for k in ("sigma_a", "fleck", "emission_E", "source_E"):
        assert _rel(ph[k], oph[k]) <= 1e-13, k
    assert _rel(ph["total_E"], oph["total_E"]) <= 1e-13
    assert rep["transported"] == orep["transported"]
    assert rep["census_size"] == orep["census_size"]
    for k in ("emission_E", "source_E", "pre_census_E", "post_census_E", "exit_E"):
        assert _rel(rep[k], orep[k]) <= 1e-12 or rep[k] == orep[k], k
