"""Independent pins of the oracle: checks fixed by mathematics, not by the
oracle's own text.

Each test is chosen so that a plausible slip in the restatement fails it:

* effective scattering (src/history_based_transport.h:60-100) in an
  infinite medium, where every history's path length is its census life:
  scatters along a path form a Poisson process of rate
  sigma_eff = (1 - f) sigma_a + sigma_s; a fraction (1 - f) sigma_a /
  sigma_eff of them re-sample the group; the weight decays as
  exp(-sigma_a f L) whatever the scatters; the tallies hold the lost weight
  and lost weight / (sigma_a f); every draw is accounted for;
* hand-traced histories and source photons: the exact draws of RNG(seed,
  stream) (pinned by the reference header's known answers) turned into
  positions, directions and weights by the published sampling rules
  (src/sampling_functions.h:22-122, src/source.h:27-62), so draw order and
  the mu / phi / sin / cos assignments are fixed;
* the group walk (src/sampling_functions.h:125-138) as inversion of its CDF;
* the step physics (src/mesh.h:237-281, 327-423) against the textbook
  implicit Monte Carlo forms of Fleck & Cummings (J. Comput. Phys. 8, 1971)
  on a deck whose opacities depend on temperature and whose temperatures
  differ between regions;
* the census (src/post_process_functions.h:32-59) on a hand-built list;
* the replicated rank share (src/mesh.h:287-315): what n ranks hold
  together against the one-rank energies.
"""
import math
import os

import numpy as np
import pytest

import oracle

A, C = oracle.A_RAD, oracle.C_LIGHT
PI = 3.1415926535897932384626433832795  # Constants::pi (src/constants.h)
DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _one_cell(sigma_a, sigma_s, fleck, half=1.0e6, bc=oracle.REFLECT):
    """One huge cell: faces out of reach of any history here."""
    G = len(np.atleast_1d(sigma_a))
    nodes = np.array([[-half, half, -half, half, -half, half]])
    return oracle.Cells(nodes, np.full((1, 6), bc, np.int32), np.zeros((1, 6), np.uint32),
                        np.reshape(sigma_a, (1, G)), np.reshape(sigma_s, (1, G)), [fleck])


def _photons(n, life, seed=4242, stream0=1000, group=0, direction=(0.6, 0.48, 0.64)):
    p = np.zeros(n, dtype=oracle.PHOTON_DTYPE)
    p["cell"] = 0
    p["group"] = group
    p["pos"] = 0.0
    p["angle"] = direction
    p["E"] = 1.0
    p["E0"] = 1.0
    p["life_dx"] = life
    p["descriptor"][:, 0] = oracle.PASS
    p["rng"][:, 0] = 0
    p["rng"][:, 1] = seed << 32
    p["rng"][:, 2] = stream0 + np.arange(n, dtype=np.uint64)
    return p


def _transport_each(cells, photons):
    """Transport photon by photon: per-history segment counts."""
    segs = np.zeros(len(photons), dtype=np.int64)
    a_tot, t_tot = np.zeros(cells.n_cells), np.zeros(cells.n_cells)
    for i in range(len(photons)):
        one = photons[i:i + 1].copy()
        a, t, s = oracle.transport(cells, one)
        photons[i] = one[0]
        segs[i] = s
        a_tot += a
        t_tot += t
    return segs, a_tot, t_tot


# ---------------------------------------------------------------------------
# effective scattering in an infinite medium
# ---------------------------------------------------------------------------
def test_effective_scatter_rate_resampling_and_decay():
    sa, ss, f, L = 10.0, 2.0, 0.4, 1.0
    sig_eff = (1.0 - f) * sa + ss  # 8 per cm; the wrong forms give 12 or 6
    p_resample = (1.0 - f) * sa / sig_eff  # 0.75; a flipped test gives 0.25
    cells = _one_cell(sa, ss, f)
    N = 6000
    ph = _photons(N, L)
    ctr0 = ph["rng"][:, 0].copy()
    segs, a_tot, t_tot = _transport_each(cells, ph)
    # no face is reached, so every segment but the last ends in a scatter
    assert np.all(ph["descriptor"][:, 0] == oracle.CENSUS)
    S = segs - 1
    # draws: one distance per segment, mu + phi + test per scatter, one
    # group draw per re-sampling scatter
    D = (ph["rng"][:, 0] - ctr0).astype(np.int64)
    R = D - segs - 3 * S
    assert np.all((R >= 0) & (R <= S))
    # scatters along a path of length L: Poisson(sig_eff L)
    m, v = S.mean(), S.var()
    assert abs(m - sig_eff * L) < 4.0 * math.sqrt(sig_eff * L / N)
    assert abs(v / m - 1.0) < 5.0 * math.sqrt(2.0 / N)
    # each scatter re-samples the group with probability (1 - f) sa / sig_eff
    frac = R.sum() / S.sum()
    assert abs(frac - p_resample) < 4.0 * math.sqrt(p_resample * (1 - p_resample) / S.sum())
    # implicit capture: weight exp(-sa f L) whatever the path; census life 0
    assert np.allclose(ph["E"], math.exp(-sa * f * L), rtol=1e-13, atol=0)
    assert np.all(ph["life_dx"] == 0.0)
    lost = (1.0 - ph["E"]).sum()
    assert abs(a_tot[0] - lost) <= 1e-12 * lost
    assert abs(t_tot[0] - lost / (sa * f)) <= 1e-12 * lost / (sa * f)
    # displacement never exceeds the path length, equals it without a scatter
    r = np.linalg.norm(ph["pos"], axis=1)
    assert np.all(r <= L * (1 + 1e-12))
    assert np.allclose(r[S == 0], L, rtol=1e-12)


def test_scattered_directions_are_isotropic():
    # after at least one scatter the final direction is the last scatter's
    # isotropic draw: <u> = <v> = <w> = 0, <u^2> = 1/3, <u v> = 0, and the
    # pairing of sin(theta) with cos(phi) / sin(phi) gives <u w> = <v w> = 0
    cells = _one_cell(1.0, 3.0, 1.0)  # sig_eff = 3: P(no scatter in L = 2) = e^-6
    ph = _photons(8000, 2.0)
    segs, _, _ = _transport_each(cells, ph)
    d = ph["angle"][segs > 1]
    n = len(d)
    assert n > 7900
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    se = 4.0 / math.sqrt(n)
    assert np.all(np.abs(d.mean(axis=0)) < se * math.sqrt(1 / 3))
    assert np.all(np.abs((d ** 2).mean(axis=0) - 1 / 3) < se * math.sqrt(4 / 45))
    for i, j in ((0, 1), (0, 2), (1, 2)):
        assert abs((d[:, i] * d[:, j]).mean()) < se * math.sqrt(1 / 15)


def test_kill_deposits_the_whole_weight():
    # sa f L = 6 > ln 100: the photon falls below the cutoff and is killed,
    # the rest of its weight deposited: the tally holds all of E0
    cells = _one_cell(6.0, 0.0, 1.0)
    ph = _photons(200, 1.0)
    segs, a_tot, _ = _transport_each(cells, ph)
    assert np.all(ph["descriptor"][:, 0] == oracle.KILLED)
    assert np.all(segs == 1)  # no scatter (sig_eff = 0): one flight, killed at its end
    assert abs(a_tot[0] - 200.0) <= 1e-12 * 200.0


def _trace_single_scatter(p0, sa, ss, f, G_sa=None):
    """Expected end state of a history with exactly one scatter, from the
    photon's draws: distance (flight 1), mu, phi, re-emission test, the group
    draw only when the test re-samples, distance (flight 2 must reach the
    census).  Returns None when the draws give another history."""
    seed = int(p0["rng"][1]) >> 32
    stream, ctr = int(p0["rng"][2]), int(p0["rng"][0])
    u = oracle.rng_draws(seed, stream, 7, ctr0=ctr)
    sig = (1.0 - f) * sa + ss
    L = float(p0["life_dx"])
    d1 = -math.log(u[0]) / sig
    if not d1 < L:
        return None
    mu = u[1] * 2.0 - 1.0
    phi = u[2] * 2.0 * PI
    st = math.sqrt(1.0 - mu * mu)
    om1 = np.array([st * math.cos(phi), st * math.sin(phi), mu])
    k = 4
    if u[3] > ss / sig:  # re-sampled: the group walk takes one draw
        k = 5
    d2 = -math.log(u[k]) / sig
    life1 = L - d1
    if not d2 > life1:
        return None
    pos = np.array(p0["pos"], dtype=float) + np.array(p0["angle"], dtype=float) * d1
    pos = pos + om1 * life1
    E = float(p0["E"])
    E = E - E * (1.0 - math.exp(-sa * f * d1))
    E = E - E * (1.0 - math.exp(-sa * f * life1))
    return dict(ctr=ctr + k + 1, pos=pos, angle=om1, E=E)


def test_single_scatter_histories_hand_traced():
    sa, ss, f, L = 2.0, 0.5, 0.5, 0.7  # sig_eff L = 1.05: a third scatter once
    cells = _one_cell(sa, ss, f)
    ph = _photons(1500, L)
    start = ph.copy()
    segs, _, _ = _transport_each(cells, ph)
    n_checked = 0
    for i in np.nonzero(segs == 2)[0]:
        exp = _trace_single_scatter(start[i], sa, ss, f)
        assert exp is not None, i  # two segments are one scatter and the census
        assert int(ph["rng"][i, 0]) == exp["ctr"]
        assert np.allclose(ph["angle"][i], exp["angle"], rtol=0, atol=1e-15)
        assert np.allclose(ph["pos"][i], exp["pos"], rtol=0, atol=1e-14)
        assert abs(ph["E"][i] - exp["E"]) <= 1e-14 * exp["E"]
        n_checked += 1
    assert n_checked > 400


# ---------------------------------------------------------------------------
# group walk: inversion of the CDF of op_a(g) / (op_a(0) G)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sa", [[1.0, 1.0, 1.0, 1.0], [2.0, 1.0, 3.0, 0.5, 1.5],
                                [1.0, 0.2, 0.2]])
def test_group_walk_inverts_its_cdf(sa):
    sa = np.array(sa)
    G = len(sa)
    cdf = np.cumsum(sa / (sa[0] * G))
    hits = np.zeros(G, dtype=np.int64)
    for stream in range(3000):
        u = oracle.rng_draws(99, stream, 1)[0]
        g = oracle.sample_emission_group(sa, 99, stream)[0]
        # the first group whose cumulative probability reaches u; past the
        # last, the last group
        want = min(int(np.searchsorted(cdf, u, side="left")), G - 1)
        if np.min(np.abs(cdf - u)) > 1e-12:
            assert g == want, (stream, u, g, want)
        hits[g] += 1
    p = np.diff(np.concatenate([[0.0], np.minimum(cdf, 1.0)]))
    p[-1] += max(0.0, 1.0 - cdf[-1])
    assert np.all(np.abs(hits / hits.sum() - p) < 4.0 * np.sqrt(p * (1 - p) / 3000) + 1e-12)


# ---------------------------------------------------------------------------
# source photons, hand traced (src/source.h:27-62, 84-131)
# ---------------------------------------------------------------------------
def test_source_photons_hand_traced():
    deck = oracle.Deck.from_xml(os.path.join(DECKS, "marshak_slab_100.xml"))
    sim = oracle.Sim(deck)
    sim.prepare()
    cells, ph, clk = sim.cells(), sim.physics(), sim.clock()
    dt = clk["dt"]
    out = oracle.make_photons(cells, ph["emission_E"], ph["source_E"], ph["source_face"],
                              20_000, ph["total_E"], deck.seed, 3, 0, dt)
    seen = {1: 0, 2: 0}
    for p in out[:: max(1, len(out) // 400)]:
        c = int(p["cell"])
        nd = cells.nodes[c]
        stream = int(p["rng"][2])
        u = oracle.rng_draws(deck.seed, stream, 7)
        if p["source_type"] == 2:  # emission: position 3, mu, phi, life, group
            pos = [nd[0] + u[0] * (nd[1] - nd[0]), nd[2] + u[1] * (nd[3] - nd[2]),
                   nd[4] + u[2] * (nd[5] - nd[4])]
            mu, phi = u[3] * 2.0 - 1.0, u[4] * 2.0 * PI
            st = math.sqrt(1.0 - mu * mu)
            ang = [st * math.cos(phi), st * math.sin(phi), mu]
            life, grp, n_draws = u[5] * C * dt, math.floor(u[6] * 1), 7
        else:  # face source on x = 0 (face 0): position 2, theta, phi, life, group
            assert p["source_type"] == 1 and ph["source_face"][c] == 0
            pos = [nd[0], nd[2] + u[0] * (nd[3] - nd[2]), nd[4] + u[1] * (nd[5] - nd[4])]
            theta, phi = math.acos(math.sqrt(u[2])), u[3] * 2.0 * PI
            ang = [math.cos(theta), math.sin(theta) * math.sin(phi),
                   math.sin(theta) * math.cos(phi)]
            life, grp, n_draws = u[4] * C * dt, math.floor(u[5] * 1), 6
        assert np.array_equal(p["pos"], pos)
        assert np.allclose(p["angle"], ang, rtol=0, atol=1e-15)
        assert p["life_dx"] == life and p["group"] == grp
        assert int(p["rng"][0]) == n_draws
        # streams of step 3, rank 0: 1e13 * 3 + photon index
        assert stream // 10**13 == 3
        seen[int(p["source_type"])] += 1
    assert seen[1] > 0 and seen[2] > 0


# ---------------------------------------------------------------------------
# step physics against the textbook implicit Monte Carlo forms
# ---------------------------------------------------------------------------
def _physics_deck():
    regions = [
        dict(ID=1, cv=0.3, rho=2.0, opA=1.0, opB=5.0, opC=3.0, opS=0.7, T_e=0.5, T_r=0.4),
        dict(ID=2, cv=0.1, rho=1.0, opA=10.0, opB=2.0, opC=-1.5, opS=0.0, T_e=1.2, T_r=0.9),
    ]
    return oracle.Deck(
        t_start=0.0, t_stop=0.006, dt_start=0.002, t_mult=1.5, dt_max=0.01,
        n_photons=20_000, seed=31, n_groups=1,
        divs=[[(0.0, 0.3, 3), (0.3, 1.0, 5)], [(0.0, 0.2, 2)], [(0.0, 0.5, 1)]],
        region_map={(0, 0, 0): 1, (1, 0, 0): 2}, regions=regions,
        bc=[oracle.SOURCE, oracle.VACUUM, oracle.REFLECT, oracle.REFLECT, oracle.VACUUM,
            oracle.VACUUM],
        T_source=0.8)


def _textbook(nodes, reg, deck, T, T_r, dt, step):
    """Fleck & Cummings: beta = 4 a T^3 / (rho cV), f = 1 / (1 + beta c dt
    sigma_a), emission f sigma_a a c T^4 V dt, a Planckian face's inflow
    a c T_s^4 / 4 per area and time, radiation energy a T_r^4 V."""
    R = deck.regions
    rho = np.array([R[r]["rho"] for r in reg])
    cv = np.array([R[r]["cv"] for r in reg])
    sa = np.array([R[r]["opA"] + R[r]["opB"] * T[i] ** R[r]["opC"] for i, r in enumerate(reg)])
    ss = np.array([R[r]["opS"] for r in reg])
    V = (nodes[:, 1] - nodes[:, 0]) * (nodes[:, 3] - nodes[:, 2]) * (nodes[:, 5] - nodes[:, 4])
    beta = 4.0 * A * T ** 3 / (rho * cv)
    f = 1.0 / (1.0 + beta * C * dt * sa)
    em = f * sa * A * C * T ** 4 * V * dt
    area_x = (nodes[:, 3] - nodes[:, 2]) * (nodes[:, 5] - nodes[:, 4])
    src = np.where(nodes[:, 0] == 0.0, 0.25 * A * C * deck.T_source ** 4 * area_x * dt, 0.0)
    cen = A * T_r ** 4 * V if step == 1 else np.zeros_like(V)
    return dict(sigma_a=sa, sigma_s=ss, fleck=f, emission_E=em, source_E=src, census_E=cen,
                V=V, rho=rho, cv=cv)


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_step_physics_textbook_forms_over_three_steps():
    deck = _physics_deck()
    sim = oracle.Sim(deck)
    nodes, _, _, reg = sim.mesh()
    T = np.array([deck.regions[r]["T_e"] for r in reg])
    T_r = np.array([deck.regions[r]["T_r"] for r in reg])
    assert len(set(T)) == 2 and len(nodes) == 16
    for step in (1, 2, 3):
        clk = sim.clock()
        sim.prepare()
        ph = sim.physics()
        tb = _textbook(nodes, reg, deck, T, T_r, clk["dt"], step)
        for k in ("sigma_a", "sigma_s"):
            assert _rel(ph[k][:, 0], tb[k]) <= 1e-13, (step, k)
        for k in ("fleck", "emission_E", "source_E", "census_E"):
            assert _rel(ph[k], tb[k]) <= 1e-13, (step, k)
        assert abs(ph["total_E"] - (tb["emission_E"] + tb["source_E"]
                                   + tb["census_E"]).sum()) <= 1e-13 * ph["total_E"]
        rep = sim.step()
        a, t = sim.tallies()
        T_new, T_r_diag = sim.temperature()
        # material energy balance: rho cV V (T_new - T) = absorbed - emitted
        lhs = tb["rho"] * tb["cv"] * tb["V"] * (T_new - T)
        assert np.allclose(lhs, a - ph["emission_E"], rtol=1e-12,
                           atol=1e-12 * np.abs(a - ph["emission_E"]).max())
        # track-length energy density: track_E / (c dt) per volume = a T_r^4
        assert np.allclose(A * T_r_diag ** 4 * tb["V"], t / (C * clk["dt"]), rtol=1e-12)
        scale = rep["pre_census_E"] + rep["emission_E"] + rep["source_E"]
        assert abs(rep["rad_conservation"]) < 1e-13 * scale
        T = T_new


# ---------------------------------------------------------------------------
# census (post_process_photons)
# ---------------------------------------------------------------------------
def test_census_keeps_input_order_and_sets_the_next_life():
    desc = [oracle.CENSUS, oracle.EXIT, oracle.KILLED, oracle.CENSUS, oracle.PASS,
            oracle.CENSUS, oracle.EXIT, oracle.CENSUS]
    p = np.zeros(len(desc), dtype=oracle.PHOTON_DTYPE)
    p["descriptor"][:, 0] = desc
    p["E"] = [0.1, 0.25, 0.0, 0.3, 0.7, 1e-17, 0.5, 0.2]
    p["cell"] = np.arange(len(desc))[::-1]
    p["life_dx"] = 123.0
    next_dt = 0.0123
    census, st = oracle.post_process(next_dt, p)
    idx = [0, 3, 5, 7]
    assert list(census["cell"]) == list(p["cell"][idx])
    assert np.all(census["life_dx"] == C * next_dt)
    assert st["census_E"] == ((0.1 + 0.3) + 1e-17) + 0.2
    assert st["exit_E"] == 0.25 + 0.5
    assert (st["n_census"], st["n_exit"], st["n_killed"]) == (4, 2, 1)


# ---------------------------------------------------------------------------
# replicated rank share
# ---------------------------------------------------------------------------
def _deck_small_cube(n_photons):
    d = oracle.Deck.from_xml(os.path.join(DECKS, "cube_40.xml"))
    d.divs = [[(0.0, 0.4, 8)], [(0.0, 0.4, 8)], [(0.0, 0.4, 8)]]
    d.n_photons = n_photons
    return d


@pytest.mark.parametrize("deck_of,n_ranks", [
    (lambda: oracle.Deck.from_xml(os.path.join(DECKS, "marshak_slab_100.xml")), 2),
    (lambda: oracle.Deck.from_xml(os.path.join(DECKS, "marshak_slab_100.xml")), 3),
    (lambda: _deck_small_cube(300), 4),   # 0.6 photons per cell: every cell reassigned
    (lambda: _deck_small_cube(200_000), 3),  # no cell reassigned
], ids=["marshak-2", "marshak-3", "cube-sparse-4", "cube-dense-3"])
def test_rank_share_sums_to_one_rank(deck_of, n_ranks):
    one = oracle.Sim(deck_of())
    one.prepare()
    p1 = one.physics()
    n_user = deck_of().n_photons
    per = []
    for r in range(n_ranks):
        s = oracle.Sim(deck_of())
        s.set_rank(r, n_ranks)
        s.prepare()
        per.append(s.physics())
    ncell = len(p1["fleck"])
    owner = np.arange(ncell) % n_ranks
    E1_tot = (p1["emission_E"] + p1["source_E"] + p1["census_E"]).sum()
    n_small = 0
    for k in ("emission_E", "source_E", "census_E"):
        stack = np.array([q[k] for q in per])  # [rank, cell]
        # every cell's energy, over the ranks, is its one-rank energy
        assert np.allclose(stack.sum(axis=0), p1[k], rtol=1e-14, atol=0), k
        # an energy that the all-rank source gives no photon of its own sits
        # whole on rank cell % n; any other is shared evenly
        small = (p1[k] > 0) & (np.floor(n_user * p1[k] / E1_tot) == 0)
        far = np.abs(n_user * p1[k] / E1_tot - np.round(n_user * p1[k] / E1_tot)) > 1e-9
        for c in np.nonzero(far & (p1[k] > 0))[0]:
            if small[c]:
                assert np.allclose(stack[owner[c], c], p1[k][c], rtol=1e-14)
                assert np.all(np.delete(stack[:, c], owner[c]) == 0.0)
            else:
                assert np.allclose(stack[:, c], p1[k][c] / n_ranks, rtol=1e-14)
        n_small += int(small.sum())
    # the fluid physics is every rank's own
    for q in per:
        assert np.array_equal(q["fleck"], p1["fleck"])
        assert np.array_equal(q["sigma_a"], p1["sigma_a"])
        # E_tot of the photon counts: all ranks together, the one-rank total
        assert abs(q["total_E"] - p1["total_E"]) <= 1e-14 * p1["total_E"]
    if "sparse" in str(n_ranks) or n_user == 300:
        assert n_small == ncell


def test_rank_streams_and_counts():
    d = oracle.Deck.from_xml(os.path.join(DECKS, "marshak_slab_100.xml"))
    n_ranks = 2
    counts = []
    for r in range(n_ranks):
        s = oracle.Sim(d)
        s.set_rank(r, n_ranks)
        s.prepare()
        cells, ph, clk = s.cells(), s.physics(), s.clock()
        out = oracle.make_photons(cells, ph["emission_E"], ph["source_E"], ph["source_face"],
                                  d.n_photons, ph["total_E"], d.seed, 1, r, clk["dt"])
        # streams 1e13 cycle + n_user rank + i (src/source.h:221-222)
        assert np.array_equal(out["rng"][:, 2].astype(np.int64),
                              10**13 + d.n_photons * r + np.arange(len(out)))
        counts.append(len(out))
        # the rank's energy, carried by its photons
        E = out["E"].sum()
        assert abs(E - (ph["emission_E"] + ph["source_E"]).sum()) <= 1e-12 * E
    # the face source is shared (n_user/2 photons each); the 100 small
    # emission cells are split between the ranks, 50 each
    assert abs(counts[0] - counts[1]) <= 2
