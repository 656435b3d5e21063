"""Tally form choice and the per-wave LDS hash tally.

Auto mode (tally_mode 0) on a mesh too large for block-private LDS tallies,
with group-independent opacities, uses one global atomic per flush through
the wave ring when the source is spread over the mesh, and the per-wave hash
table when it is concentrated (a face source: 1,600 of 64,000 cells).  Both
forms must give the oracle's events exactly and its tallies to summation
order (1e-12 relative).
"""
import numpy as np
import pytest

import oracle
from branson_amd import capi, problems

pytestmark = pytest.mark.gpu

RING, LDS, TWO, HASH = 1, 2, 3, 4


def _box(kind):
    n = 40
    faces = (capi.SOURCE if kind == "face" else capi.VACUUM,) + (capi.VACUUM,) * 5
    nodes, bc, nxt = problems.box_mesh(n, n, n, 0.05, 0.05, 0.05, faces)
    m = len(nodes)
    rng = np.random.default_rng(17)
    cells = oracle.Cells(nodes, bc, nxt, rng.uniform(5.0, 40.0, m).reshape(-1, 1),
                         np.full((m, 1), 1.0), rng.uniform(0.3, 0.9, m))
    em, src = np.zeros(m), np.zeros(m)
    face = np.full(m, -1, dtype=np.int8)
    if kind == "face":
        lit = nodes[:, 0] == 0.0  # the x = 0 layer
        src[lit] = 1.0
        face[lit] = 0
    else:
        em[:] = 1.0
    return cells, em, src, face


def _photons(cells, em, src, face, n_user=400_000):
    return oracle.make_photons(cells, em, src, face, n_user, em.sum() + src.sum(), 31, 1, 0,
                               0.001)


def _check(got, ref, a_g, t_g, a_o, t_o, st, s_o):
    assert np.array_equal(got["descriptor"][:, 0], ref["descriptor"][:, 0])
    assert np.array_equal(got["cell"], ref["cell"])
    assert np.array_equal(got["rng"], ref["rng"])
    assert st["n_segments"] == s_o
    assert np.allclose(got["E"], ref["E"], rtol=1e-12, atol=0)
    assert np.allclose(a_g, a_o, rtol=1e-12, atol=1e-300)
    assert np.allclose(t_g, t_o, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("kind,mode,expect", [("face", 0, HASH), ("face", 1, RING),
                                              ("emission", 0, RING), ("emission", 4, HASH)])
@pytest.mark.parametrize("algo", [capi.ALGO_HISTORY, capi.ALGO_PERSISTENT])
def test_tally_forms_match_oracle(kind, mode, expect, algo):
    cells, em, src, face = _box(kind)
    photons = _photons(cells, em, src, face)
    ref = photons.copy()
    a_o, t_o, s_o = oracle.transport(cells, ref, 8)
    got = photons.copy()
    with capi.Engine(0, algorithm=algo, tally_mode=mode) as e:
        e.set_mesh_cells(cells)
        a_g, t_g, st = e.transport_photons(got)
    assert st["tally"] == expect
    _check(got, ref, a_g, t_g, a_o, t_o, st, s_o)


@pytest.mark.parametrize("kind,expect", [("face", HASH), ("emission", RING)])
def test_auto_choice_from_source_energies(kind, expect):
    """The device-resident step decides from the per-cell source energies."""
    cells, em, src, face = _box(kind)
    n_user = 300_000
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        e.set_mesh_cells(cells)
        e.source(em, src, None, face, n_user, em.sum() + src.sum(), 31, 1, 0, 0.001)
        photons = e.photons()
        a_g, t_g, st = e.transport(0.001)
    assert st["tally"] == expect
    ref = photons.copy()
    a_o, t_o, s_o = oracle.transport(cells, ref, 8)
    assert st["n_segments"] == s_o
    assert np.allclose(a_g, a_o, rtol=1e-12, atol=1e-300)
    assert np.allclose(t_g, t_o, rtol=1e-12, atol=1e-300)


def test_hash_tally_fp32_events_equal_fp64():
    cells, em, src, face = _box("face")
    photons = _photons(cells, em, src, face, 200_000)
    out = {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK):
        got = photons.copy()
        with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT, precision=prec, fp32_tol=1e-6) as e:
            e.set_mesh_cells(cells)
            a, t, st = e.transport_photons(got)
        assert st["tally"] == HASH
        out[prec] = (got, a, st)
    g64, a64, _ = out[capi.PRECISION_FP64]
    g32, a32, st32 = out[capi.PRECISION_FP32_TRACK]
    assert np.array_equal(g32["descriptor"][:, 0], g64["descriptor"][:, 0])
    assert np.array_equal(g32["rng"], g64["rng"])
    m = a64 > 0
    assert np.max(np.abs(a32[m] - a64[m]) / a64[m]) <= 1e-6
