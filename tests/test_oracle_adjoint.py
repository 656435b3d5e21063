"""Pins the oracle's differential-operator (likelihood-ratio) tally, the CPU
reference the GPU adjoint is checked against.

What fixes it, independently of its own formulas:
  * a pure absorber (f = 1, sigma_s = 0) samples no flights: the tally is a
    deterministic function of the opacities, so the estimator must equal the
    exact derivative, checked by central differences with a tiny step;
  * a closed reflecting box without the cutoff: every history absorbs
    E0 (1 - exp(-sigma_a f life)) whatever its path, so the expected total
    derivative is E0 f life exp(-sigma_a f life) (sigma_a direction) and
    E0 sigma_a life exp(-sigma_a f life) (Fleck direction) -- the
    likelihood-ratio terms must average to zero;
  * a 6^3 vacuum cube with a random opacity field: central differences with
    common random numbers (h = 1e-4 relative) agree within statistical error
    (both estimate the derivative of the expected tally);
  * the photons' events are those of the plain transport.
"""
import math

import numpy as np
import pytest

import oracle

C_LIGHT = 299.792458


def _box_cells(nx, width, sigma_a, sigma_s, fleck, bc_kind):
    n = nx ** 3
    nodes = np.zeros((n, 6))
    bc = np.zeros((n, 6), dtype=np.int32)
    nxt = np.zeros((n, 6), dtype=np.uint32)
    for k in range(nx):
        for j in range(nx):
            for i in range(nx):
                c = i + nx * (j + nx * k)
                nodes[c] = [i * width, (i + 1) * width, j * width, (j + 1) * width,
                            k * width, (k + 1) * width]
                for face, (di, dj, dk) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0),
                                                     (0, 1, 0), (0, 0, -1), (0, 0, 1)]):
                    a, b, e = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < nx and 0 <= e < nx:
                        bc[c, face] = oracle.ELEMENT
                        nxt[c, face] = a + nx * (b + nx * e)
                    else:
                        bc[c, face] = bc_kind
                        nxt[c, face] = c
    sa = np.broadcast_to(np.asarray(sigma_a, dtype=float), (n,)).reshape(n, 1)
    ss = np.broadcast_to(np.asarray(sigma_s, dtype=float), (n,)).reshape(n, 1)
    f = np.broadcast_to(np.asarray(fleck, dtype=float), (n,))
    return oracle.Cells(nodes, bc, nxt, sa, ss, f)


def _with(cells, sigma_a=None, fleck=None):
    sa = cells.sigma_a if sigma_a is None else np.asarray(sigma_a).reshape(-1, 1)
    f = cells.fleck if fleck is None else np.asarray(fleck)
    return oracle.Cells(cells.nodes, cells.bc, cells.next, sa, cells.sigma_s, f)


def _emit(cells, n_user, seed, dt):
    em = np.ones(cells.n_cells)
    return oracle.make_photons(cells, em, np.zeros(cells.n_cells),
                               np.full(cells.n_cells, -1, dtype=np.int8), n_user,
                               em.sum(), seed, 1, 0, dt)


def test_dos_events_match_plain_transport():
    cells = _box_cells(4, 0.05, 5.0, 1.0, 0.5, oracle.VACUUM)
    ph = _emit(cells, 20000, 7, 0.001)
    p1, p2 = ph.copy(), ph.copy()
    a1, _, s1 = oracle.transport(cells, p1, 1)
    v = np.ones(cells.n_cells)
    a2, _, s2 = oracle.dos_transport(cells, p2, v, v, 4)
    assert s1 == s2
    assert np.array_equal(p1, p2)
    assert np.allclose(a1, a2, rtol=1e-12, atol=0)


def test_dos_pure_absorber_equals_exact_derivative():
    # f = 1, sigma_s = 0: no flight is sampled and the tally is a smooth
    # function of sigma_a away from the (measure-zero) cutoff crossings
    rng = np.random.default_rng(3)
    cells = _box_cells(3, 0.05, rng.uniform(2.0, 8.0, 27), 0.0, 1.0, oracle.VACUUM)
    ph = _emit(cells, 30000, 11, 0.001)
    v = rng.normal(size=cells.n_cells)
    _, dabs, _ = oracle.dos_transport(cells, ph.copy(), v, np.zeros(cells.n_cells), 4)
    h = 1e-6
    sa = cells.sigma_a[:, 0]
    ap, _, _ = oracle.transport(_with(cells, sigma_a=sa + h * v), ph.copy(), 4)
    am, _, _ = oracle.transport(_with(cells, sigma_a=sa - h * v), ph.copy(), 4)
    fd = (ap - am) / (2 * h)
    assert np.allclose(dabs, fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())


@pytest.mark.parametrize("direction", ["sigma_a", "fleck"])
def test_dos_reflecting_box_total_derivative_closed_form(direction):
    sa, ss, f, dt = 5.0, 1.0, 0.5, 0.001  # sa f c dt = 0.75: no cutoff kills
    cells = _box_cells(2, 0.05, sa, ss, f, oracle.REFLECT)
    n = cells.n_cells
    ph = _emit(cells, 40000, 21, dt)
    life = ph["life_dx"].copy()
    E0 = ph["E"].copy()
    one, zero = np.ones(n), np.zeros(n)
    va, vf = (one, zero) if direction == "sigma_a" else (zero, one)
    # per-photon totals: run the photons in 20 batches for a standard error
    batches = np.array_split(np.arange(len(ph)), 20)
    est = []
    for b in batches:
        _, dabs, _ = oracle.dos_transport(cells, ph[b].copy(), va, vf, 1)
        est.append(dabs.sum())
    est = np.array(est)
    att = np.exp(-sa * f * life)
    exact_per = E0 * (f if direction == "sigma_a" else sa) * life * att
    exact = np.array([exact_per[b].sum() for b in batches])
    diff = est - exact
    se = diff.std(ddof=1) / math.sqrt(len(diff))
    assert abs(diff.mean()) < 4 * se + 1e-12 * abs(exact.mean())
    # the estimator is not trivially exact: the likelihood-ratio terms move it
    assert se > 0


def test_dos_matches_central_differences_with_common_random_numbers():
    rng = np.random.default_rng(17)
    nx = 6
    n = nx ** 3
    sa = np.exp(rng.normal(np.log(5.0), 0.5, n))
    cells = _box_cells(nx, 0.05, sa, 1.0, rng.uniform(0.3, 0.8, n), oracle.VACUUM)
    ph = _emit(cells, 60000, 5, 0.001)
    va = rng.uniform(0.5, 1.5, n) * sa   # one-signed: a resolvable derivative
    vf = -rng.uniform(0.05, 0.15, n)
    w = rng.uniform(0.5, 1.5, n)  # a fixed linear functional of the tallies
    h = 1e-4
    batches = np.array_split(np.arange(len(ph)), 20)
    dos, fd = [], []
    for b in batches:
        _, dabs, _ = oracle.dos_transport(cells, ph[b].copy(), va, vf, 1)
        dos.append(w @ dabs)
        ap, _, _ = oracle.transport(_with(cells, sa + h * va, cells.fleck + h * vf),
                                    ph[b].copy(), 1)
        am, _, _ = oracle.transport(_with(cells, sa - h * va, cells.fleck - h * vf),
                                    ph[b].copy(), 1)
        fd.append(w @ (ap - am) / (2 * h))
    dos, fd = np.array(dos), np.array(fd)
    d = dos - fd
    se = d.std(ddof=1) / math.sqrt(len(d))
    print("dos", dos.mean(), "fd", fd.mean(), "se(diff)", se)
    assert abs(d.mean()) < 3 * se
    # the derivative is resolved: well away from zero at this sample size
    assert abs(dos.mean()) > 5 * dos.std(ddof=1) / math.sqrt(len(dos))
