"""Many independent problems in one engine (bh_set_instances), the batched
workload of many problems of uneven size.

  * bh_source_instances equals each instance's own make_photons (its seed,
    photon count and energy, streams numbered within the instance), photon
    for photon, concatenated with the cell offsets;
  * one FP64 transport over all instances (one work-item per history,
    persistent waves, event-split passes) equals the oracle's transport of
    the same photons on the concatenated mesh;
  * FP32 tracking over the batch: every event as FP64, tallies within the
    FP32 tolerance;
  * new photons made inside the transport kernel, each from its instance's
    record and stream, equal the source kernel's (census bit for bit);
  * a whole batched step with the physics on the device
    (bh_step_resident_instances) on 8 instances spanning the configuration's
    range of sizes (50 to 32768 cells, 1e4 to 1e6 photons) equals the
    oracle's per-instance source and transport fed the device's per-cell
    physics: photon counts, census and exits per instance, census order and
    per-cell tallies to 1e-12;
  * the full configuration -- 256 instances, three FP64 steps -- keeps every
    instance's energy balance to round-off, accounts for every photon, and
    keeps each instance's seed on its census photons; steps 2 and 3 run as
    one HIP graph each and give the photon counts of the eager steps.
"""
import numpy as np
import pytest

import oracle
from branson_amd import capi, problems
from branson_amd.batched import BatchedRun

pytestmark = pytest.mark.gpu


def _small_batch():
    return problems.batched_instances(n_instances=12, seed=4, cells_range=(20, 800),
                                      photons_range=(1000, 20000))


def _host_energies(b):
    """Emission energies and per-instance totals from the problems' own step
    physics on the host (inputs for the engine's instance entry points)."""
    em = problems.emission_energy(b.nodes, b.cat("sigma_a"), b.cat("fleck"), b.cat("T"),
                                  b.parts[0].dt)
    return em, np.bincount(b.instance_of_cell(), weights=em, minlength=b.n_instances)


def _engine_with_batch(b, **kw):
    e = capi.Engine(0, **kw)
    e.set_mesh(b.nodes, b.bc, b.next, b.cat("sigma_a"), b.cat("sigma_s"), b.cat("fleck"))
    e.set_instances(b.first_cell, b.seeds)
    return e


def _oracle_cells(b, sigma_a, sigma_s, fleck):
    return oracle.Cells(b.nodes, b.bc, b.next, np.reshape(sigma_a, (-1, 1)),
                        np.reshape(sigma_s, (-1, 1)), fleck)


def _oracle_source(b, sigma_a, sigma_s, fleck, emission, total_E, cycle, dt):
    out = []
    for k, p in enumerate(b.parts):
        lo, hi = int(b.first_cell[k]), int(b.first_cell[k + 1])
        cells = oracle.Cells(p.nodes, p.bc, p.next, np.reshape(sigma_a[lo:hi], (-1, 1)),
                             np.reshape(sigma_s[lo:hi], (-1, 1)), fleck[lo:hi])
        e = emission[lo:hi]
        ph = oracle.make_photons(cells, e, np.zeros(len(e)), np.full(len(e), -1, dtype=np.int8),
                                 int(b.n_photons[k]), float(total_E[k]), int(b.seeds[k]), cycle,
                                 0, dt)
        ph["cell"] += lo
        out.append(ph)
    return np.concatenate(out)


def test_batched_source_equals_per_instance_sources():
    b = _small_batch()
    em, total_E = _host_energies(b)
    n, dt = b.n_cells, b.parts[0].dt
    with _engine_with_batch(b) as e:
        e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8), b.n_photons,
                           total_E, 1, 0, dt)
        got = e.photons()
    ref = _oracle_source(b, b.cat("sigma_a"), b.cat("sigma_s"), b.cat("fleck"), em, total_E, 1,
                         dt)
    assert len(got) == len(ref)
    for f in ("cell", "group", "source_type", "rng", "pos", "E", "E0", "life_dx"):
        assert np.array_equal(got[f], ref[f]), f
    assert np.allclose(got["angle"], ref["angle"], rtol=0, atol=4e-16)
    # every instance draws from its own seed
    seeds = got["rng"][:, 1] >> np.uint64(32)
    assert np.array_equal(seeds, b.seeds[b.instance_of_cell()[got["cell"]]])


@pytest.mark.parametrize("algo", [capi.ALGO_HISTORY, capi.ALGO_PERSISTENT, capi.ALGO_EVENT])
def test_batched_transport_fp64_matches_oracle(algo):
    b = _small_batch()
    em, total_E = _host_energies(b)
    n, dt = b.n_cells, b.parts[0].dt
    with _engine_with_batch(b, algorithm=algo) as e:
        e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8), b.n_photons,
                           total_E, 1, 0, dt)
        photons = e.photons()
        got = photons.copy()
        a_g, t_g, st = e.transport_photons(got)
    ref = photons.copy()
    a_o, t_o, s_o = oracle.transport(
        _oracle_cells(b, b.cat("sigma_a"), b.cat("sigma_s"), b.cat("fleck")), ref, 8)
    assert np.array_equal(got["descriptor"][:, 0], ref["descriptor"][:, 0])
    assert np.array_equal(got["cell"], ref["cell"])
    assert np.array_equal(got["rng"], ref["rng"])
    assert st["n_segments"] == s_o
    assert np.allclose(a_g, a_o, rtol=1e-12, atol=1e-300)
    assert np.allclose(t_g, t_o, rtol=1e-12, atol=1e-300)


def test_batched_instance_lattice_records_equal_general_records():
    """Every problem of the batch is a lattice, so the persistent FP64 kernel
    tracks on per-instance 32-byte records and axes (stats lattice 2).  The
    same arithmetic runs on the same numbers as on the general records
    (general_mesh), so every photon agrees bit for bit and the tallies to
    summation order."""
    b = _small_batch()
    em, total_E = _host_energies(b)
    n, dt = b.n_cells, b.parts[0].dt
    out = {}
    for general in (False, True):
        with _engine_with_batch(b, algorithm=capi.ALGO_PERSISTENT, general_mesh=general) as e:
            e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8),
                               b.n_photons, total_E, 1, 0, dt)
            ph = e.photons()
            a, t, st = e.transport_photons(ph)
            out[general] = (ph, a, t, st)
    (pl, al, tl, sl), (pg, ag, tg, sg) = out[False], out[True]
    assert sl["lattice"] == 2 and sg["lattice"] == 0
    for f in ("descriptor", "cell", "group", "rng", "pos", "angle", "E", "life_dx"):
        assert np.array_equal(pl[f], pg[f]), f
    for k in ("n_segments", "n_census", "n_exit", "n_killed", "n_flushes"):
        assert sl[k] == sg[k], k
    assert np.allclose(al, ag, rtol=1e-12, atol=0)
    assert np.allclose(tl, tg, rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK])
def test_batched_photons_made_in_transport_equal_source_kernel(precision):
    """New photons made by the persistent kernel where it transports them
    (source_mode SOURCE_IN_TRANSPORT), each from its own instance's record and
    stream, against the source kernel's: the census bit for bit and in order,
    the photon counts and census energy identical, tallies to 1e-12."""
    b = _small_batch()
    em, total_E = _host_energies(b)
    n, dt = b.n_cells, b.parts[0].dt
    out = {}
    for mode in (capi.SOURCE_KERNEL, capi.SOURCE_IN_TRANSPORT):
        with _engine_with_batch(b, algorithm=capi.ALGO_PERSISTENT, precision=precision,
                                source_mode=mode) as e:
            e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8),
                               b.n_photons, total_E, 1, 0, dt)
            a, t, st = e.transport(dt)
            out[mode] = (a, t, st, e.census())
    (ak, tk, sk, ck), (ag, tg, sg, cg) = out[capi.SOURCE_KERNEL], out[capi.SOURCE_IN_TRANSPORT]
    for k in ("n_transported", "n_census", "n_exit", "n_killed", "n_segments", "census_E"):
        assert sk[k] == sg[k], k
    assert len(ck) == len(cg) and np.array_equal(ck.view(np.uint8), cg.view(np.uint8))
    assert np.allclose(ak, ag, rtol=1e-12, atol=0)
    assert np.allclose(tk, tg, rtol=1e-12, atol=0)


def test_batched_fp32_tracking_events_match_fp64():
    b = _small_batch()
    em, total_E = _host_energies(b)
    n, dt = b.n_cells, b.parts[0].dt
    out = {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK):
        with _engine_with_batch(b, algorithm=capi.ALGO_PERSISTENT, precision=prec,
                                fp32_tol=1e-6) as e:
            e.source_instances(em, np.zeros(n), None, np.full(n, -1, dtype=np.int8),
                               b.n_photons, total_E, 1, 0, dt)
            ph = e.photons()
            a, t, st = e.transport_photons(ph)
            out[prec] = (ph, a, st)
    p64, a64, s64 = out[capi.PRECISION_FP64]
    p32, a32, s32 = out[capi.PRECISION_FP32_TRACK]
    assert np.array_equal(p32["descriptor"][:, 0], p64["descriptor"][:, 0])
    assert np.array_equal(p32["cell"], p64["cell"])
    assert s32["n_segments"] == s64["n_segments"]
    m = a64 != 0
    assert np.max(np.abs(a32[m] - a64[m]) / a64[m]) <= 1e-6


def _spanning_batch():
    """8 instances from the configuration's range: cells 50 ... 32768 and
    photons 1e4 ... 1e6, each geometric, the largest problem with the
    fewest photons and the other way round in between."""
    rng = np.random.default_rng(17)
    cells = np.round(np.geomspace(50, 32768, 8)).astype(int)
    photons = np.round(np.geomspace(1e4, 1e6, 8)).astype(int)[[0, 5, 2, 7, 4, 1, 6, 3]]
    parts = []
    for k, nc in enumerate(cells):
        if k % 2 == 0:
            kinds = (capi.VACUUM, capi.VACUUM) + (capi.REFLECT,) * 4
            nodes, bc, nxt = problems.box_mesh(int(nc), 1, 1, 0.02, 0.02, 0.02, kinds)
        else:
            m = max(2, int(round(nc ** (1 / 3))))
            nodes, bc, nxt = problems.box_mesh(m, m, m, 0.02, 0.02, 0.02, capi.VACUUM)
        n = len(nodes)
        sa = 10.0 * np.exp(rng.normal(0.0, 1.0, n))
        parts.append(problems.CellProblem(nodes, bc, nxt, sa, rng.uniform(0, 2, n),
                                          rng.uniform(0.01, 1.0), 0.01))
    return problems.Batch(parts, photons, np.arange(8) * 7 + 3)


def test_batched_resident_step_matches_oracle_slice():
    b = _spanning_batch()
    dt = b.parts[0].dt
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        run = BatchedRun(b, e)
        r = run.step()
        ph = e.physics()
        a_g, t_g = e.tallies()
        cen = e.census()
    per = r["instances"]
    sa, ss, fl, em = ph["sigma_a"][:, 0], ph["sigma_s"][:, 0], ph["fleck"], ph["emission_E"]
    # the device's per-cell physics is the textbook form of each instance
    T = b.cat("T")
    cv = np.concatenate([np.full(p.n_cells, p.cv) for p in b.parts])
    rho = np.concatenate([np.full(p.n_cells, p.rho) for p in b.parts])
    assert np.allclose(fl, problems.fleck_factor(b.cat("sigma_a"), T, dt, cv, rho), rtol=1e-13)
    assert np.allclose(em, problems.emission_energy(b.nodes, b.cat("sigma_a"), fl, T, dt),
                       rtol=1e-13)
    inst = b.instance_of_cell()
    assert np.allclose(per["total_E"], np.bincount(inst, weights=em, minlength=8), rtol=1e-13)
    # the oracle: each instance's make_photons with the device's energies and
    # E_tot, then one transport over the concatenated mesh
    photons = _oracle_source(b, sa, ss, fl, em, per["total_E"], 1, dt)
    counts = np.bincount(inst[photons["cell"]], minlength=8)
    assert np.array_equal(per["n_new"], counts)
    a_o, t_o, _ = oracle.transport(_oracle_cells(b, sa, ss, fl), photons, 8)
    census, st = oracle.post_process(dt, photons)
    assert r["n_census"] == st["n_census"] and r["n_exit"] == st["n_exit"]
    assert r["n_killed"] == st["n_killed"]
    assert np.array_equal(cen["cell"], census["cell"])
    ex = photons["descriptor"][:, 0] == oracle.EXIT
    exit_k = np.bincount(inst[photons["cell"][ex]], weights=photons["E"][ex], minlength=8)
    post_k = np.bincount(inst[census["cell"]], weights=census["E"], minlength=8)

    def rel(x, y):
        m = y != 0
        return float(np.max(np.abs(x[m] - y[m]) / np.abs(y[m]))) if m.any() else 0.0
    errs = {"photon_E": rel(cen["E"], census["E"]), "abs_E": rel(a_g, a_o),
            "track_E": rel(t_g, t_o), "exit_E": rel(per["exit_E"], exit_k),
            "census_E": rel(per["post_census_E"], post_k)}
    print("batched slice, max relative error against the oracle:", errs, flush=True)
    # per-cell absorbed and census energy: the 1e-12 bar.  A single census
    # photon's weight carries the last-ulp difference of every exp on its
    # path (device ocml against host glibc), and implicit capture amplifies
    # it by 1 / (survival fraction) per segment: over the hundreds of
    # segments a photon makes in these lognormal slabs that reaches 1e-11.
    assert errs["abs_E"] <= 1e-12 and errs["track_E"] <= 1e-12
    assert errs["census_E"] <= 1e-12 and errs["exit_E"] <= 1e-12
    assert errs["photon_E"] <= 1e-10
    assert np.allclose(per["absorbed_E"], np.bincount(inst, weights=a_o, minlength=8),
                       rtol=1e-12)


def test_batched_config_256_instances_three_steps():
    b = problems.batched_instances()  # 256 instances, 50-32768 cells, 1e4-1e6 photons
    assert b.n_instances == 256
    print(f"batch: {b.n_cells} cells, {int(b.n_photons.sum())} photons/step", flush=True)
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT,
                     reserve_photons=int(b.n_photons.sum()) * 3) as e:
        run = BatchedRun(b, e)
        for step in (1, 2, 3):
            r = run.step()
            print(f"step {step}: {r['n_transported']} histories, census {r['n_census']}, "
                  f"graph {r['graph']}, balance {r['max_instance_balance']:.2e}", flush=True)
            assert r["graph"] == (1 if step > 1 else 0)
            assert r["max_instance_balance"] <= 1e-12
            assert r["n_census"] + r["n_exit"] + r["n_killed"] == r["n_transported"]
            per = r["instances"]
            assert np.all(per["n_new"] > 0)
            assert int(per["n_new"].sum()) <= r["n_transported"]
        cen = e.census()
    seeds = cen["rng"][:, 1] >> np.uint64(32)
    assert np.array_equal(seeds, b.seeds[b.instance_of_cell()[cen["cell"]]])


def test_batched_graph_steps_equal_eager_steps():
    b = problems.batched_instances(n_instances=24, seed=9, cells_range=(50, 4000),
                                   photons_range=(1e4, 1e5))
    out = {}
    for algo in (capi.ALGO_PERSISTENT, capi.ALGO_HISTORY):
        with capi.Engine(0, algorithm=algo) as e:
            run = BatchedRun(b, e)
            out[algo] = [run.step() for _ in range(3)]
    g, h = out[capi.ALGO_PERSISTENT], out[capi.ALGO_HISTORY]
    assert [r["graph"] for r in g] == [0, 1, 1] and [r["graph"] for r in h] == [0, 0, 0]
    for a, c in zip(g, h):
        assert a["n_transported"] == c["n_transported"]
        assert a["n_census"] == c["n_census"]
        assert np.array_equal(a["instances"]["n_new"], c["instances"]["n_new"])
        for k in ("absorbed_E", "post_census_E", "exit_E", "post_mat_E"):
            x, y = a["instances"][k], c["instances"][k]
            assert np.allclose(x, y, rtol=1e-9, atol=1e-300), k
