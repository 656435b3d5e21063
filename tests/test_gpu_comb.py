"""Census population control on the device (bh_comb_census) against the
oracle's comb_photons (src/census_functions.h:48-93).

Survivors must be the oracle's photons in the oracle's order (cells, RNG
words, positions exact); their energies agree to the rounding of comb_E
(exactly when the census energy was summed serially); every cell's census
energy is conserved.  A crafted census exercises the two count rules (a
cell's last photon kept when the others are combed away; a lone photon
kept), and a step transported after the comb conserves energy.
"""
import os

import numpy as np
import pytest

import oracle
from branson_amd import capi, problems

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _per_cell(ph, n):
    return np.bincount(ph["cell"], weights=ph["E"], minlength=n)


def _check(got, ref, rep, before, n_cells):
    assert len(got) == len(ref) == rep["n_after"]
    for f in ("cell", "group", "rng", "pos", "angle", "life_dx", "descriptor"):
        assert np.array_equal(got[f], ref[f]), f
    tol = 0.0 if rep["comb_serial"] else 8 * np.spacing(rep["comb_E"])
    assert np.all(np.abs(got["E"] - ref["E"]) <= tol)
    assert np.allclose(_per_cell(got, n_cells), _per_cell(before, n_cells), rtol=1e-12,
                       atol=8 * np.spacing(rep["comb_E"]))


@pytest.mark.parametrize("max_census", [50_000, 200_000])
def test_comb_equals_oracle_on_a_transported_census(max_census):
    d = oracle.Deck.from_xml(os.path.join(GOLDEN, "decks", "cube_40.xml"))
    sim = oracle.Sim(d)
    sim.prepare()
    cells, ph, clk = sim.cells(), sim.physics(), sim.clock()
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        e.set_mesh_cells(cells)
        e.source(ph["emission_E"], ph["source_E"], None, ph["source_face"], 400_000,
                 ph["total_E"], d.seed, 1, 0, clk["dt"])
        e.transport(clk["next_dt"])
        before = e.census()
        assert len(before) > max_census // 2
        rep = e.comb_census(max_census, 1234, 99, 17)
        got = e.census()
    ref = oracle.comb_photons(before, cells.n_cells, max_census, 1234, 99, 17)
    assert rep["n_before"] == len(before)
    _check(got, ref, rep, before, cells.n_cells)
    assert abs(rep["census_E_after"] - before["E"].sum()) <= 1e-12 * before["E"].sum()


def _box_cells(n=8):
    nodes, bc, nxt = problems.box_mesh(n, n, n, 0.1, 0.1, 0.1, capi.VACUUM)
    m = len(nodes)
    return oracle.Cells(nodes, bc, nxt, np.full((m, 1), 2.0), np.full((m, 1), 1.0),
                        np.full(m, 0.5))


def test_comb_count_rules_on_a_crafted_census():
    cells = _box_cells()
    rng = np.random.default_rng(3)
    cell_of = rng.integers(1, cells.n_cells, 3000).astype(np.uint32)
    # cell 0: four photons of negligible energy (combed away but the last);
    # a lone photon in the last cell
    cell_of[[10, 500, 1700, 2999]] = 0
    cell_of[1234] = cells.n_cells - 1
    cell_of[cell_of == cells.n_cells - 1] = 1
    cell_of[1234] = cells.n_cells - 1
    E = np.where(cell_of == 0, 1e-12, rng.lognormal(0.0, 0.5, len(cell_of)))
    census = np.zeros(len(cell_of), dtype=capi.PHOTON_DTYPE)
    census["cell"] = cell_of
    nd = cells.nodes[cell_of]
    census["pos"] = 0.5 * (nd[:, 0::2] + nd[:, 1::2])
    census["angle"] = [0.0, 0.0, 1.0]
    census["E"] = E
    census["E0"] = E
    census["life_dx"] = 0.3
    census["descriptor"][:, 0] = capi.CENSUS
    census["rng"][:, 0] = 7
    census["rng"][:, 1] = 55 << 32
    census["rng"][:, 2] = np.arange(len(census)) + 10**6
    with capi.Engine(0) as e:
        e.set_mesh_cells(cells)
        e.census_import(census)
        rep = e.comb_census(400, 55, 3, 0)
        got = e.census()
    ref = oracle.comb_photons(census, cells.n_cells, 400, 55, 3, 0)
    _check(got, ref, rep, census, cells.n_cells)
    c0 = got[got["cell"] == 0]
    assert len(c0) == 1 and c0["rng"][0, 2] == 2999 + 10**6
    lone = got[got["cell"] == cells.n_cells - 1]
    assert len(lone) == 1 and lone["rng"][0, 2] == 1234 + 10**6


def test_transport_after_comb_conserves_energy():
    d = oracle.Deck.from_xml(os.path.join(GOLDEN, "decks", "marshak_slab_100.xml"))
    with capi.Driver(os.path.join(GOLDEN, "decks", "marshak_slab_100.xml"),
                     algorithm=capi.ALGO_PERSISTENT) as drv:
        drv.step()
        L = capi.lib()
        h = L.bh_driver_engine(drv._d)
        rep = capi.CombReport()
        n = capi.C.c_uint64()
        assert L.bh_census_size(h, capi.C.byref(n)) == 0
        assert L.bh_comb_census(h, max(1, n.value // 4), 5, 6, 0, capi.C.byref(rep)) == 0
        assert rep.n_after < rep.n_before
        assert abs(rep.census_E_after - rep.census_E_before) <= 1e-12 * rep.census_E_before
        # the driver's next step starts from the combed census
        r = drv.step()
        assert r["census_size"] > 0
        scale = max(r["post_mat_E"], r["absorbed_E"] + r["exit_E"] + r["post_census_E"])
        # pre_census_E is the driver's record of the census before the comb,
        # equal to the combed energy to round-off
        assert abs(r["rad_conservation"]) < 1e-11 * scale
    assert d.n_photons > 0
