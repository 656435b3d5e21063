"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  This package is the correctness oracle and the timed
CPU baseline: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check results
against (or, in the benchmark, to time the CPU).
The product package (``branson_amd``) never imports it.

The C restatement in ``branson_oracle.c`` follows the reference file:line cited
at each function.  The deck reader below restates the XML schema handling of
src/input.h:105-490 (missing numeric fields read as 0, as pugixml's
``text().as_double()`` does).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_KAT = os.path.join(HERE, "_ref", "threefry_kat")

# src/constants.h:26-28
REFLECT, VACUUM, ELEMENT, SOURCE, PROCESSOR = range(5)
EXIT, PASS, CENSUS, SCATTER, KILLED, BOUND = range(6)
C_LIGHT = 299.792458
A_RAD = 0.01372

# 120-byte AOS Photon of src/photon.h:172-181
PHOTON_DTYPE = np.dtype(
    [
        ("cell", "<u4"),
        ("group", "<u4"),
        ("source_type", "<u4"),
        ("descriptor", "u1", (4,)),
        ("pos", "<f8", (3,)),
        ("angle", "<f8", (3,)),
        ("E", "<f8"),
        ("E0", "<f8"),
        ("life_dx", "<f8"),
        ("rng", "<u8", (4,)),
    ]
)
assert PHOTON_DTYPE.itemsize == 120


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    return LIB_PATH


class _Cells(C.Structure):
    _fields_ = [
        ("n_cells", C.c_uint32),
        ("n_groups", C.c_uint32),
        ("nodes", C.c_void_p),
        ("bc", C.c_void_p),
        ("next", C.c_void_p),
        ("sigma_a", C.c_void_p),
        ("sigma_s", C.c_void_p),
        ("fleck", C.c_void_p),
    ]


class _CensusStats(C.Structure):
    _fields_ = [
        ("census_E", C.c_double),
        ("exit_E", C.c_double),
        ("n_census", C.c_uint64),
        ("n_exit", C.c_uint64),
        ("n_killed", C.c_uint64),
    ]


class _Deck(C.Structure):
    _fields_ = [
        ("t_start", C.c_double),
        ("t_stop", C.c_double),
        ("dt_start", C.c_double),
        ("t_mult", C.c_double),
        ("dt_max", C.c_double),
        ("n_photons", C.c_uint64),
        ("seed", C.c_uint32),
        ("n_groups", C.c_uint32),
        ("n_div", C.c_uint32 * 3),
        ("div_start", C.c_void_p * 3),
        ("div_end", C.c_void_p * 3),
        ("div_cells", C.c_void_p * 3),
        ("div_region", C.c_void_p),
        ("n_regions", C.c_uint32),
        ("cv", C.c_void_p),
        ("rho", C.c_void_p),
        ("opA", C.c_void_p),
        ("opB", C.c_void_p),
        ("opC", C.c_void_p),
        ("opS", C.c_void_p),
        ("T_e", C.c_void_p),
        ("T_r", C.c_void_p),
        ("bc", C.c_int32 * 6),
        ("T_source", C.c_double),
    ]


class _StepReport(C.Structure):
    _fields_ = [
        ("step", C.c_uint32),
        ("time", C.c_double),
        ("dt", C.c_double),
        ("next_dt", C.c_double),
        ("transported", C.c_uint64),
        ("census_size", C.c_uint64),
        ("segments", C.c_uint64),
        ("emission_E", C.c_double),
        ("source_E", C.c_double),
        ("absorbed_E", C.c_double),
        ("exit_E", C.c_double),
        ("pre_census_E", C.c_double),
        ("post_census_E", C.c_double),
        ("pre_mat_E", C.c_double),
        ("post_mat_E", C.c_double),
        ("rad_conservation", C.c_double),
        ("mat_conservation", C.c_double),
        ("transport_seconds", C.c_double),
        ("source_seconds", C.c_double),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P, U64, U32, D, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_int
        L.orc_threefry2x64_20.argtypes = [P, P, P]
        L.orc_u01.argtypes = [U64]
        L.orc_u01.restype = D
        L.orc_rng_draws.argtypes = [U32, U64, U64, U32, P]
        L.orc_sample_position_in_cell.argtypes = [P, U32, U64, P, P]
        L.orc_sample_position_on_face.argtypes = [P, I, U32, U64, P, P]
        L.orc_sample_uniform_angle.argtypes = [U32, U64, P, P]
        L.orc_sample_source_angle.argtypes = [I, U32, U64, P, P]
        L.orc_sample_emission_group.argtypes = [P, U32, U32, U64, P]
        L.orc_sample_emission_group.restype = I
        L.orc_distance_to_boundary.argtypes = [P, P, P, P]
        L.orc_distance_to_boundary.restype = D
        L.orc_source_counts.argtypes = [U32, P, P, U64, D, P]
        L.orc_source_counts.restype = U64
        L.orc_make_photons.argtypes = [P, P, P, P, U64, D, U32, U32, U32, D, P, U64]
        L.orc_make_photons.restype = U64
        L.orc_make_initial_census.argtypes = [P, P, U64, D, U32, U32, D, P, U64]
        L.orc_make_initial_census.restype = U64
        L.orc_transport.argtypes = [P, P, U64, P, P, I]
        L.orc_transport.restype = U64
        L.orc_dos_transport.argtypes = [P, P, U64, P, P, P, P, I]
        L.orc_dos_transport.restype = U64
        L.orc_post_process.argtypes = [D, P, U64, P, U64, P]
        L.orc_post_process.restype = U64
        L.orc_comb_photons.argtypes = [P, U64, C.c_uint32, C.c_int64, C.c_uint32, U64, U64]
        L.orc_comb_photons.restype = U64
        L.orc_sim_create.argtypes = [P]
        L.orc_sim_create.restype = P
        L.orc_sim_destroy.argtypes = [P]
        L.orc_sim_set_rank.argtypes = [P, U32, U32]
        L.orc_sim_n_cells.argtypes = [P]
        L.orc_sim_n_cells.restype = U32
        L.orc_sim_finished.argtypes = [P]
        L.orc_sim_finished.restype = I
        L.orc_sim_step.argtypes = [P, I, P]
        L.orc_sim_step.restype = I
        L.orc_sim_prepare.argtypes = [P]
        L.orc_sim_prepare.restype = I
        L.orc_sim_get_mesh.argtypes = [P, P, P, P, P]
        L.orc_sim_get_physics.argtypes = [P, P, P, P, P, P, P, P, P]
        L.orc_sim_get_tallies.argtypes = [P, P, P]
        L.orc_sim_get_temperature.argtypes = [P, P, P]
        L.orc_sim_census_size.argtypes = [P]
        L.orc_sim_census_size.restype = U64
        L.orc_sim_get_census.argtypes = [P, P]
        L.orc_sim_clock.argtypes = [P, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


# --------------------------------------------------------------------------
# RNG and samplers
# --------------------------------------------------------------------------
def threefry2x64_20(ctr0: int, ctr1: int, key0: int, key1: int) -> tuple[int, int]:
    c = np.array([ctr0, ctr1], dtype=np.uint64)
    k = np.array([key0, key1], dtype=np.uint64)
    o = np.zeros(2, dtype=np.uint64)
    lib().orc_threefry2x64_20(_ptr(c), _ptr(k), _ptr(o))
    return int(o[0]), int(o[1])


def u01(x: int) -> float:
    return lib().orc_u01(x)


def rng_draws(seed: int, stream: int, n: int, ctr0: int = 0) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().orc_rng_draws(seed, stream, ctr0, n, _ptr(out))
    return out


def sample_position_in_cell(nodes, seed, stream):
    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    pos = np.zeros(3)
    ctr = np.zeros(1, dtype=np.uint64)
    lib().orc_sample_position_in_cell(_ptr(nd), seed, stream, _ptr(pos), _ptr(ctr))
    return pos, int(ctr[0])


def sample_position_on_face(nodes, face, seed, stream):
    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    pos = np.zeros(3)
    ctr = np.zeros(1, dtype=np.uint64)
    lib().orc_sample_position_on_face(_ptr(nd), face, seed, stream, _ptr(pos), _ptr(ctr))
    return pos, int(ctr[0])


def sample_uniform_angle(seed, stream):
    ang = np.zeros(3)
    ctr = np.zeros(1, dtype=np.uint64)
    lib().orc_sample_uniform_angle(seed, stream, _ptr(ang), _ptr(ctr))
    return ang, int(ctr[0])


def sample_source_angle(face, seed, stream):
    ang = np.zeros(3)
    ctr = np.zeros(1, dtype=np.uint64)
    lib().orc_sample_source_angle(face, seed, stream, _ptr(ang), _ptr(ctr))
    return ang, int(ctr[0])


def sample_emission_group(sigma_a_cell, seed, stream):
    sa = np.ascontiguousarray(sigma_a_cell, dtype=np.float64)
    ctr = np.zeros(1, dtype=np.uint64)
    g = lib().orc_sample_emission_group(_ptr(sa), len(sa), seed, stream, _ptr(ctr))
    return g, int(ctr[0])


def distance_to_boundary(nodes, pos, angle):
    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    a = np.ascontiguousarray(angle, dtype=np.float64)
    sc = np.zeros(1, dtype=np.uint32)
    d = lib().orc_distance_to_boundary(_ptr(nd), _ptr(p), _ptr(a), _ptr(sc))
    return d, int(sc[0])


# --------------------------------------------------------------------------
# Cells, source, transport, census
# --------------------------------------------------------------------------
class Cells:
    """Cell-major read-only cell arrays (src/cell.h:319-336)."""

    def __init__(self, nodes, bc, next_, sigma_a, sigma_s, fleck):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 6)
        n = self.nodes.shape[0]
        self.bc = np.ascontiguousarray(bc, dtype=np.int32).reshape(n, 6)
        self.next = np.ascontiguousarray(next_, dtype=np.uint32).reshape(n, 6)
        self.sigma_a = np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(n, -1)
        self.sigma_s = np.ascontiguousarray(sigma_s, dtype=np.float64).reshape(n, -1)
        self.fleck = np.ascontiguousarray(fleck, dtype=np.float64).reshape(n)
        self.n_cells = n
        self.n_groups = self.sigma_a.shape[1]
        self._c = _Cells(
            n, self.n_groups, _ptr(self.nodes), _ptr(self.bc), _ptr(self.next),
            _ptr(self.sigma_a), _ptr(self.sigma_s), _ptr(self.fleck),
        )

    @property
    def ref(self):
        return C.byref(self._c)


def source_counts(emission_E, source_E, n_user, total_E):
    e = np.ascontiguousarray(emission_E, dtype=np.float64)
    s = np.ascontiguousarray(source_E, dtype=np.float64)
    counts = np.zeros(2 * len(e), dtype=np.uint32)
    tot = lib().orc_source_counts(len(e), _ptr(e), _ptr(s), n_user, total_E, _ptr(counts))
    return int(tot), counts.reshape(-1, 2)


def make_photons(cells: Cells, emission_E, source_E, source_face, n_user, total_E,
                 seed, cycle, rank, dt) -> np.ndarray:
    e = np.ascontiguousarray(emission_E, dtype=np.float64)
    s = np.ascontiguousarray(source_E, dtype=np.float64)
    f = np.ascontiguousarray(source_face, dtype=np.int8)
    L = lib()
    n = L.orc_make_photons(cells.ref, _ptr(e), _ptr(s), _ptr(f), n_user, total_E,
                           seed, cycle, rank, dt, None, 0)
    out = np.zeros(n, dtype=PHOTON_DTYPE)
    L.orc_make_photons(cells.ref, _ptr(e), _ptr(s), _ptr(f), n_user, total_E,
                       seed, cycle, rank, dt, _ptr(out), n)
    return out


def make_initial_census(cells: Cells, census_E, n_user, total_E, seed, rank, dt):
    e = np.ascontiguousarray(census_E, dtype=np.float64)
    L = lib()
    n = L.orc_make_initial_census(cells.ref, _ptr(e), n_user, total_E, seed, rank,
                                  dt, None, 0)
    out = np.zeros(n, dtype=PHOTON_DTYPE)
    L.orc_make_initial_census(cells.ref, _ptr(e), n_user, total_E, seed, rank, dt,
                              _ptr(out), n)
    return out


def transport(cells: Cells, photons: np.ndarray, n_threads: int = 1):
    """Transport in place; returns (abs_E, track_E, n_segments)."""
    assert photons.dtype == PHOTON_DTYPE and photons.flags.c_contiguous
    abs_E = np.zeros(cells.n_cells)
    track_E = np.zeros(cells.n_cells)
    segs = lib().orc_transport(cells.ref, _ptr(photons), len(photons), _ptr(abs_E),
                               _ptr(track_E), n_threads)
    return abs_E, track_E, int(segs)


def dos_transport(cells: Cells, photons: np.ndarray, v_sigma_a, v_fleck, n_threads: int = 1):
    """Transport in place with the forward-mode differential-operator tally
    along (v_sigma_a, v_fleck); returns (abs_E, dabs_E, n_segments), where
    dabs_E[c] = d abs_E[c] / dt for sigma_a + t v_sigma_a, fleck + t v_fleck."""
    assert photons.dtype == PHOTON_DTYPE and photons.flags.c_contiguous
    va = np.ascontiguousarray(v_sigma_a, dtype=np.float64)
    vf = np.ascontiguousarray(v_fleck, dtype=np.float64)
    assert va.shape == (cells.n_cells,) and vf.shape == (cells.n_cells,)
    abs_E = np.zeros(cells.n_cells)
    dabs_E = np.zeros(cells.n_cells)
    segs = lib().orc_dos_transport(cells.ref, _ptr(photons), len(photons), _ptr(va), _ptr(vf),
                                   _ptr(abs_E), _ptr(dabs_E), n_threads)
    return abs_E, dabs_E, int(segs)


def post_process(next_dt: float, photons: np.ndarray):
    census = np.zeros(len(photons), dtype=PHOTON_DTYPE)
    st = _CensusStats()
    lib().orc_post_process(next_dt, _ptr(photons), len(photons), _ptr(census),
                           len(photons), C.byref(st))
    census = census[: st.n_census].copy()
    return census, {
        "census_E": st.census_E, "exit_E": st.exit_E, "n_census": st.n_census,
        "n_exit": st.n_exit, "n_killed": st.n_killed,
    }


def comb_photons(photons: np.ndarray, n_cells: int, max_census: int, seed: int, stream: int,
                 ctr0: int = 0) -> np.ndarray:
    """comb_photons (src/census_functions.h:48-93) on a census in order, with
    RNG(seed, stream) from counter ctr0.  Returns the survivors in order."""
    ph = np.ascontiguousarray(photons, dtype=PHOTON_DTYPE).copy()
    k = lib().orc_comb_photons(_ptr(ph), len(ph), n_cells, max_census, seed, stream, ctr0)
    return ph[:k].copy()


# --------------------------------------------------------------------------
# Decks (XML schema of src/input.h:105-490) and the step driver
# --------------------------------------------------------------------------
_BC = {"REFLECT": REFLECT, "VACUUM": VACUUM, "SOURCE": SOURCE}
# face order X_NEG X_POS Y_NEG Y_POS Z_NEG Z_POS (src/input.h:345-415)
_BC_TAGS = ["bc_left", "bc_right", "bc_down", "bc_up", "bc_bottom", "bc_top"]


def _num(node, tag, conv=float, default=0):
    if node is None:
        return default
    c = node.find(tag)
    if c is None or c.text is None or not c.text.strip():
        return default
    return conv(c.text.strip())


class Deck:
    """Plain description of a Branson deck: times, divisions, regions, bcs."""

    def __init__(self, **kw):
        self.t_start = kw.get("t_start", 0.0)
        self.t_stop = kw["t_stop"]
        self.dt_start = kw["dt_start"]
        self.t_mult = kw.get("t_mult", 1.0)
        self.dt_max = kw.get("dt_max", 1.0)
        self.n_photons = int(kw["n_photons"])
        self.seed = int(kw["seed"])
        self.n_groups = int(kw.get("n_groups", 1))
        self.divs = kw["divs"]  # 3 lists of (start, end, n_cells)
        self.region_map = kw["region_map"]  # {(ix, iy, iz): region_ID}
        self.regions = kw["regions"]  # list of dicts with ID, cv, rho, ...
        self.bc = kw["bc"]  # 6 ints
        self.T_source = kw.get("T_source", 0.0)

    @classmethod
    def from_xml(cls, path: str, n_groups: int = 1) -> "Deck":
        root = ET.parse(path).getroot()
        com = root.find("common")
        sp = root.find("spatial")
        bnd = root.find("boundary")
        regs = root.find("regions")
        divs = [[], [], []]
        rmap = {}
        for ch in sp:
            for a, ax in enumerate("xyz"):
                if ch.tag == f"{ax}_division":
                    divs[a].append((_num(ch, f"{ax}_start"), _num(ch, f"{ax}_end"),
                                    _num(ch, f"n_{ax}_cells", int)))
            if ch.tag == "region_map":
                rmap[(_num(ch, "x_div_ID", int), _num(ch, "y_div_ID", int),
                      _num(ch, "z_div_ID", int))] = _num(ch, "region_ID", int)
        bc = []
        for t in _BC_TAGS:
            txt = (bnd.findtext(t) or "").strip()
            if txt not in _BC:
                raise ValueError(f"boundary type not recognized: {t}={txt!r}")
            bc.append(_BC[txt])
        regions = []
        for r in regs.findall("region"):
            regions.append(dict(
                ID=_num(r, "ID", int), cv=_num(r, "CV"), rho=_num(r, "density"),
                opA=_num(r, "opacA"), opB=_num(r, "opacB"), opC=_num(r, "opacC"),
                opS=_num(r, "opacS"), T_e=_num(r, "initial_T_e"),
                T_r=_num(r, "initial_T_r"),
            ))
        T_source = _num(bnd, "T_source") if SOURCE in bc else 0.0
        return cls(
            t_start=_num(com, "t_start"), t_stop=_num(com, "t_stop"),
            dt_start=_num(com, "dt_start"), t_mult=_num(com, "t_mult"),
            dt_max=_num(com, "dt_max"), n_photons=_num(com, "photons", int),
            seed=_num(com, "seed", int) & 0xFFFFFFFF, n_groups=n_groups,
            divs=divs, region_map=rmap, regions=regions, bc=bc, T_source=T_source,
        )


class Sim:
    """Replicated-mode driver of the oracle (src/replicated_driver.h:33-122)."""

    def __init__(self, deck: Deck):
        self.deck = deck
        keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        d = _Deck()
        d.t_start, d.t_stop, d.dt_start = deck.t_start, deck.t_stop, deck.dt_start
        d.t_mult, d.dt_max = deck.t_mult, deck.dt_max
        d.n_photons, d.seed, d.n_groups = deck.n_photons, deck.seed, deck.n_groups
        for a in range(3):
            d.n_div[a] = len(deck.divs[a])
            d.div_start[a] = arr([v[0] for v in deck.divs[a]], np.float64)
            d.div_end[a] = arr([v[1] for v in deck.divs[a]], np.float64)
            d.div_cells[a] = arr([v[2] for v in deck.divs[a]], np.uint32)
        id_to_idx = {r["ID"]: i for i, r in enumerate(deck.regions)}
        nx, ny, nz = (len(v) for v in deck.divs)
        if len(deck.region_map) != nx * ny * nz:
            raise ValueError("number of divisions must match the region maps")
        dr = np.zeros((nz, ny, nx), dtype=np.uint32)
        for (ix, iy, iz), rid in deck.region_map.items():
            dr[iz, iy, ix] = id_to_idx[rid]
        d.div_region = arr(dr, np.uint32)
        d.n_regions = len(deck.regions)
        for k in ("cv", "rho", "opA", "opB", "opC", "opS", "T_e", "T_r"):
            setattr(d, k, arr([r[k] for r in deck.regions], np.float64))
        for f in range(6):
            d.bc[f] = deck.bc[f]
        d.T_source = deck.T_source
        self._h = lib().orc_sim_create(C.byref(d))
        del keep
        self.n_cells = lib().orc_sim_n_cells(self._h)
        self.n_groups = deck.n_groups

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_sim_destroy(h)
            self._h = None

    def set_rank(self, rank: int, n_ranks: int):
        """Run as replicated rank `rank` of `n_ranks` (src/mesh.h:80-87,
        287-315): a 1/n_ranks share of every energy, the small-energy
        reassignment, the all-rank E_tot and the rank's stream offsets."""
        lib().orc_sim_set_rank(self._h, rank, n_ranks)

    def finished(self) -> bool:
        return bool(lib().orc_sim_finished(self._h))

    def prepare(self):
        lib().orc_sim_prepare(self._h)

    def clock(self) -> dict:
        """time, dt, next_dt and step number of the coming step."""
        t, dt, nd = C.c_double(), C.c_double(), C.c_double()
        st = C.c_uint32()
        lib().orc_sim_clock(self._h, C.byref(t), C.byref(dt), C.byref(nd), C.byref(st))
        return {"time": t.value, "dt": dt.value, "next_dt": nd.value, "step": st.value}

    def step(self, n_threads: int = 1) -> dict:
        r = _StepReport()
        if lib().orc_sim_step(self._h, n_threads, C.byref(r)) != 0:
            raise RuntimeError("simulation already finished")
        return {k: getattr(r, k) for k, _ in _StepReport._fields_}

    def run(self, n_threads: int = 1) -> list[dict]:
        reps = []
        while not self.finished():
            reps.append(self.step(n_threads))
        return reps

    def mesh(self):
        n = self.n_cells
        nodes = np.zeros((n, 6))
        bc = np.zeros((n, 6), dtype=np.int32)
        nxt = np.zeros((n, 6), dtype=np.uint32)
        reg = np.zeros(n, dtype=np.uint32)
        lib().orc_sim_get_mesh(self._h, _ptr(nodes), _ptr(bc), _ptr(nxt), _ptr(reg))
        return nodes, bc, nxt, reg

    def physics(self):
        n, G = self.n_cells, self.n_groups
        sa, ss = np.zeros((n, G)), np.zeros((n, G))
        f, em, src, cen = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        face = np.zeros(n, dtype=np.int8)
        tot = np.zeros(1)
        lib().orc_sim_get_physics(self._h, _ptr(sa), _ptr(ss), _ptr(f), _ptr(em),
                                  _ptr(src), _ptr(cen), _ptr(face), _ptr(tot))
        return dict(sigma_a=sa, sigma_s=ss, fleck=f, emission_E=em, source_E=src,
                    census_E=cen, source_face=face, total_E=float(tot[0]))

    def cells(self) -> Cells:
        nodes, bc, nxt, _ = self.mesh()
        ph = self.physics()
        return Cells(nodes, bc, nxt, ph["sigma_a"], ph["sigma_s"], ph["fleck"])

    def tallies(self):
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        lib().orc_sim_get_tallies(self._h, _ptr(a), _ptr(t))
        return a, t

    def temperature(self):
        te, tr = np.zeros(self.n_cells), np.zeros(self.n_cells)
        lib().orc_sim_get_temperature(self._h, _ptr(te), _ptr(tr))
        return te, tr

    def census(self) -> np.ndarray:
        n = lib().orc_sim_census_size(self._h)
        out = np.zeros(n, dtype=PHOTON_DTYPE)
        lib().orc_sim_get_census(self._h, _ptr(out))
        return out
