"""ctypes binding of libbranson_hip.so (include/branson_hip.h).

Thin by design: the engine and its host driver are C++/HIP; this module only
marshals numpy arrays across the C-ABI for tests, the benchmark and the
autograd wrapper.  Loading fails loudly when the library is missing -- there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbranson_hip.so")

BH_OK = 0
ERRORS = {-1: "BH_ERR_ARG", -2: "BH_ERR_DEVICE", -3: "BH_ERR_ALLOC", -4: "BH_ERR_STATE",
          -5: "BH_ERR_CAPACITY"}
REFLECT, VACUUM, ELEMENT, SOURCE, PROCESSOR = range(5)
EXIT, PASS, CENSUS, SCATTER, KILLED, BOUND = range(6)
PRECISION_FP64, PRECISION_FP32_TRACK = 0, 1
ALGO_HISTORY, ALGO_PERSISTENT, ALGO_EVENT = 0, 1, 2
SOURCE_AUTO, SOURCE_IN_TRANSPORT, SOURCE_KERNEL = 0, 1, 2

# the reference's 120-byte AOS Photon (src/photon.h:172-181)
PHOTON_DTYPE = np.dtype([
    ("cell", "<u4"), ("group", "<u4"), ("source_type", "<u4"), ("descriptor", "u1", (4,)),
    ("pos", "<f8", (3,)), ("angle", "<f8", (3,)), ("E", "<f8"), ("E0", "<f8"),
    ("life_dx", "<f8"), ("rng", "<u8", (4,)),
])
assert PHOTON_DTYPE.itemsize == 120

# every symbol include/branson_hip.h declares
EXPORTS = [
    "bh_abi_version", "bh_create", "bh_destroy", "bh_last_error", "bh_sync", "bh_set_mesh",
    "bh_set_source_energy", "bh_source", "bh_source_from_energy", "bh_set_instances",
    "bh_source_instances", "bh_transport", "bh_set_material", "bh_step_resident",
    "bh_get_temperature", "bh_get_physics", "bh_comb_census", "bh_step_resident_instances",
    "bh_step_graph", "bh_get_tallies", "bh_census_size", "bh_census_export",
    "bh_photons_export",
    "bh_census_import",
    "bh_transport_photons", "bh_transport_photons_noalloc", "bh_adjoint_photons",
    "bh_rng_draws", "bh_threefry", "bh_copy_probe",
]


class BhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("algorithm", C.c_int32), ("block_size", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("tally_mode", C.c_int32),
                ("refill_threshold", C.c_int32), ("reserve_photons", C.c_uint64),
                ("fp32_tol", C.c_double), ("batch_photons", C.c_int32),
                ("step_graph", C.c_int32), ("general_mesh", C.c_int32),
                ("source_mode", C.c_int32)]


class Cells(C.Structure):
    _fields_ = [("n_cells", C.c_uint32), ("n_groups", C.c_uint32), ("nodes", C.c_void_p),
                ("bc", C.c_void_p), ("next", C.c_void_p), ("sigma_a", C.c_void_p),
                ("sigma_s", C.c_void_p), ("fleck", C.c_void_p)]


class StepStats(C.Structure):
    _fields_ = [("census_E", C.c_double), ("exit_E", C.c_double),
                ("n_transported", C.c_uint64), ("n_census", C.c_uint64),
                ("n_exit", C.c_uint64), ("n_killed", C.c_uint64), ("n_segments", C.c_uint64),
                ("n_flushes", C.c_uint64), ("transport_ms", C.c_double),
                ("census_ms", C.c_double), ("source_ms", C.c_double),
                ("n_replayed", C.c_uint64), ("lattice", C.c_int32), ("tally", C.c_int32),
                ("replay_ms", C.c_double), ("n_allocs", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class InstanceReport(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("total_E", "emission_E", "source_E", "pre_census_E",
                                          "post_census_E", "absorbed_E", "exit_E",
                                          "pre_mat_E", "post_mat_E")] + [("n_new", C.c_uint64)]


class Material(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("opA", "opB", "opC", "opS", "cv", "rho", "T_e", "T_r",
                                          "T_s", "source_face")]


class ResidentReport(C.Structure):
    _fields_ = [("total_E", C.c_double), ("emission_E", C.c_double), ("source_E", C.c_double),
                ("pre_census_E", C.c_double), ("pre_mat_E", C.c_double),
                ("post_mat_E", C.c_double), ("absorbed_E", C.c_double), ("stats", StepStats),
                ("n_new", C.c_uint64), ("etot_serial", C.c_int32), ("graph", C.c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "stats"}
        d["stats"] = self.stats.as_dict()
        return d


class CombReport(C.Structure):
    _fields_ = [("n_before", C.c_uint64), ("n_after", C.c_uint64), ("comb_E", C.c_double),
                ("census_E_before", C.c_double), ("census_E_after", C.c_double),
                ("comb_serial", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


_lib = None
ABI_VERSION = 4  # BH_ABI_VERSION of include/branson_hip.h


def lib() -> C.CDLL:
    """Load libbranson_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m branson_amd.build`")
    L = C.CDLL(LIB_PATH)
    P, U32, U64, D, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
    L.bh_abi_version.restype = I
    L.bh_create.argtypes = [I, P, P]
    L.bh_destroy.argtypes = [P]
    L.bh_destroy.restype = None
    L.bh_last_error.argtypes = [P]
    L.bh_last_error.restype = C.c_char_p
    L.bh_sync.argtypes = [P]
    L.bh_set_mesh.argtypes = [P, P]
    L.bh_set_source_energy.argtypes = [P, P, P, P, P]
    L.bh_source.argtypes = [P, U64, D, U32, U32, U32, D, P]
    L.bh_source_from_energy.argtypes = [P, P, P, P, P, U64, D, U32, U32, U32, D, P]
    L.bh_set_instances.argtypes = [P, U32, P, P]
    L.bh_source_instances.argtypes = [P, P, P, U32, U32, D, P]
    L.bh_transport.argtypes = [P, D, P, P, P]
    L.bh_step_graph.argtypes = [P, U64, D, U32, U32, U32, D, D, P, P, P]
    L.bh_get_tallies.argtypes = [P, P, P]
    L.bh_census_size.argtypes = [P, P]
    L.bh_census_export.argtypes = [P, P, U64, P]
    L.bh_census_import.argtypes = [P, P, U64]
    L.bh_photons_export.argtypes = [P, P, U64, P]
    L.bh_transport_photons.argtypes = [P, P, U64, P, P, P]
    L.bh_transport_photons_noalloc.argtypes = [P, P, U64, P, P, P]
    L.bh_adjoint_photons.argtypes = [P, P, U64, P, P, P, P]
    L.bh_rng_draws.argtypes = [P, U32, U64, U64, U32, P]
    L.bh_threefry.argtypes = [P, P, U64, P]
    L.bh_copy_probe.argtypes = [P, U64, I, P]
    L.bh_set_material.argtypes = [P, P, D, U32, U32]
    L.bh_step_resident.argtypes = [P, U64, U32, U32, U32, D, D, P]
    L.bh_step_resident_instances.argtypes = [P, P, U32, U32, D, D, P, P]
    L.bh_get_temperature.argtypes = [P, P, P]
    L.bh_get_physics.argtypes = [P, P, P, P, P, P]
    L.bh_comb_census.argtypes = [P, C.c_int64, U32, U64, U64, P]
    for name in EXPORTS:
        if name not in ("bh_destroy", "bh_last_error", "bh_abi_version"):
            getattr(L, name).restype = I
    if L.bh_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {L.bh_abi_version()}, these bindings "
                          f"expect {ABI_VERSION}: rebuild it with `python -m branson_amd.build`")
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


class Engine:
    """One engine handle on one HIP device (not thread safe)."""

    def __init__(self, device: int = 0, precision: int = PRECISION_FP64,
                 algorithm: int = ALGO_HISTORY, tally_mode: int = 0, blocks_per_cu: int = 0,
                 refill_threshold: int = 0, reserve_photons: int = 0, fp32_tol: float = 0.0,
                 block_size: int = 0, batch_photons: int = 0, general_mesh: bool = False,
                 source_mode: int = 0):
        L = lib()
        self._h = C.c_void_p()
        self.opts = Options(precision, algorithm, block_size, blocks_per_cu, tally_mode,
                            refill_threshold, reserve_photons, fp32_tol, batch_photons, 0,
                            int(general_mesh), source_mode)
        rc = L.bh_create(device, C.byref(self.opts), C.byref(self._h))
        if rc != BH_OK:
            msg = L.bh_last_error(self._h).decode() if self._h else "bh_create failed"
            if self._h:
                L.bh_destroy(self._h)
                self._h = None
            raise BhError(rc, msg)
        self.n_cells = 0
        self.n_groups = 0
        self._n_new = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().bh_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != BH_OK:
            raise BhError(rc, lib().bh_last_error(self._h).decode())

    # -- mesh and source ---------------------------------------------------
    def set_mesh(self, nodes, bc, next_, sigma_a, sigma_s, fleck):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 6)
        n = nodes.shape[0]
        self._mesh_keep = (
            nodes,
            np.ascontiguousarray(bc, dtype=np.int32).reshape(n, 6),
            np.ascontiguousarray(next_, dtype=np.uint32).reshape(n, 6),
            np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(n, -1),
            np.ascontiguousarray(sigma_s, dtype=np.float64).reshape(n, -1),
            np.ascontiguousarray(fleck, dtype=np.float64).reshape(n),
        )
        k = self._mesh_keep
        G = k[3].shape[1]
        c = Cells(n, G, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), _p(k[5]))
        self._check(lib().bh_set_mesh(self._h, C.byref(c)))
        self.n_cells, self.n_groups = n, G

    def set_mesh_cells(self, cells):
        """From an object with nodes/bc/next/sigma_a/sigma_s/fleck arrays."""
        self.set_mesh(cells.nodes, cells.bc, cells.next, cells.sigma_a, cells.sigma_s,
                      cells.fleck)

    def source(self, emission_E, source_E, census_E, source_face, n_user, total_E, seed,
               cycle, rank, dt) -> int:
        e = np.ascontiguousarray(emission_E, dtype=np.float64)
        s = np.ascontiguousarray(source_E, dtype=np.float64)
        c = None if census_E is None else np.ascontiguousarray(census_E, dtype=np.float64)
        f = np.ascontiguousarray(source_face, dtype=np.int8)
        n_new = C.c_uint64()
        self._check(lib().bh_source_from_energy(self._h, _p(e), _p(s), _p(c), _p(f), n_user,
                                                total_E, seed, cycle, rank, dt,
                                                C.byref(n_new)))
        self._n_new = n_new.value
        return n_new.value

    def step_graph(self, n_user, total_E, seed, cycle, rank, dt, next_dt):
        """bh_step_graph: source + transport + census as one graph launch."""
        st = StepStats()
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(lib().bh_step_graph(self._h, n_user, total_E, seed, cycle, rank, dt, next_dt,
                                        _p(a), _p(t), C.byref(st)))
        self._n_new = 0
        return a, t, st.as_dict()

    def set_instances(self, first_cell, seeds):
        """Instances of the mesh: cells [first_cell[k], first_cell[k+1]) with
        their own seeds (bh_set_instances); empty first_cell clears them."""
        fc = np.ascontiguousarray(first_cell, dtype=np.uint32)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        k = max(len(fc) - 1, 0)
        if k and len(sd) != k:
            raise ValueError("one seed per instance")
        self._inst_keep = (fc, sd)
        self._check(lib().bh_set_instances(self._h, k, _p(fc) if k else None,
                                           _p(sd) if k else None))

    def source_instances(self, emission_E, source_E, census_E, source_face, n_user, total_E,
                         cycle, rank, dt) -> int:
        e = np.ascontiguousarray(emission_E, dtype=np.float64)
        s = np.ascontiguousarray(source_E, dtype=np.float64)
        c = None if census_E is None else np.ascontiguousarray(census_E, dtype=np.float64)
        f = np.ascontiguousarray(source_face, dtype=np.int8)
        self._check(lib().bh_set_source_energy(self._h, _p(e), _p(s), _p(c), _p(f)))
        nu = np.ascontiguousarray(n_user, dtype=np.uint64)
        te = np.ascontiguousarray(total_E, dtype=np.float64)
        n_new = C.c_uint64()
        self._check(lib().bh_source_instances(self._h, _p(nu), _p(te), cycle, rank, dt,
                                              C.byref(n_new)))
        self._n_new = n_new.value
        return n_new.value

    def transport(self, next_dt, want_tallies=True):
        st = StepStats()
        a = np.zeros(self.n_cells) if want_tallies else None
        t = np.zeros(self.n_cells) if want_tallies else None
        self._check(lib().bh_transport(self._h, next_dt, _p(a), _p(t), C.byref(st)))
        self._n_new = 0
        return a, t, st.as_dict()

    def tallies(self):
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(lib().bh_get_tallies(self._h, _p(a), _p(t)))
        return a, t

    # -- census ------------------------------------------------------------
    def census_size(self) -> int:
        n = C.c_uint64()
        self._check(lib().bh_census_size(self._h, C.byref(n)))
        return n.value

    def census(self) -> np.ndarray:
        n = self.census_size()
        out = np.zeros(n, dtype=PHOTON_DTYPE)
        got = C.c_uint64()
        self._check(lib().bh_census_export(self._h, _p(out), n, C.byref(got)))
        return out[: got.value]

    def photons(self) -> np.ndarray:
        """All resident photons in transport order (new, then census)."""
        got = C.c_uint64()
        n = self.census_size() + self._n_new
        out = np.zeros(n, dtype=PHOTON_DTYPE)
        self._check(lib().bh_photons_export(self._h, _p(out), n, C.byref(got)))
        return out[: got.value]

    def census_import(self, photons: np.ndarray):
        ph = np.ascontiguousarray(photons, dtype=PHOTON_DTYPE)
        self._check(lib().bh_census_import(self._h, _p(ph), len(ph)))

    def transport_photons(self, photons: np.ndarray):
        """gpu_transport_photons drop-in: transports AOS photons in place."""
        assert photons.dtype == PHOTON_DTYPE and photons.flags.c_contiguous
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        st = StepStats()
        self._check(lib().bh_transport_photons(self._h, _p(photons), len(photons), _p(a),
                                               _p(t), C.byref(st)))
        return a, t, st.as_dict()

    def transport_photons_noalloc(self, photons: np.ndarray):
        """transport_photons through the entry point that allocates no device
        memory (its buffers come from the library's static scratch array)."""
        assert photons.dtype == PHOTON_DTYPE and photons.flags.c_contiguous
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        st = StepStats()
        self._check(lib().bh_transport_photons_noalloc(self._h, _p(photons), len(photons),
                                                       _p(a), _p(t), C.byref(st)))
        return a, t, st.as_dict()

    def adjoint_photons(self, photons: np.ndarray, abs_adjoint):
        """Path-replay adjoint: (dL/dsigma_a, dL/dfleck, stats) per cell for
        photons in their pre-transport state and abs_adjoint = dL/d abs_E."""
        assert photons.dtype == PHOTON_DTYPE and photons.flags.c_contiguous
        g = np.ascontiguousarray(abs_adjoint, dtype=np.float64)
        if g.shape != (self.n_cells,):
            raise ValueError(f"abs_adjoint must have {self.n_cells} entries")
        da, df = np.zeros(self.n_cells), np.zeros(self.n_cells)
        st = StepStats()
        self._check(lib().bh_adjoint_photons(self._h, _p(photons), len(photons), _p(g), _p(da),
                                             _p(df), C.byref(st)))
        return da, df, st.as_dict()

    def comb_census(self, max_census: int, seed: int, stream: int, ctr0: int = 0) -> dict:
        """bh_comb_census: comb_photons (src/census_functions.h:48-93) on the
        resident census, draws from RNG(seed, stream) at counter ctr0."""
        rep = CombReport()
        self._check(lib().bh_comb_census(self._h, max_census, seed, stream, ctr0,
                                         C.byref(rep)))
        return rep.as_dict()

    # -- device-resident step physics -----------------------------------------
    def set_material(self, opA, opB, opC, opS, cv, rho, T_e, T_r, T_s, source_face,
                     replicated_factor=1.0, n_ranks=1, rank=0):
        """bh_set_material: per-cell opacity law, heat capacity, density,
        initial temperatures and SOURCE face, for bh_step_resident."""
        n = self.n_cells
        cols = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
                for a in (opA, opB, opC, opS, cv, rho, T_e, T_r, T_s)]
        face = np.ascontiguousarray(np.broadcast_to(np.asarray(source_face, dtype=np.int8),
                                                    (n,)))
        m = Material(*[_p(c) for c in cols], _p(face))
        self._check(lib().bh_set_material(self._h, C.byref(m), replicated_factor, n_ranks,
                                          rank))

    def step_resident(self, n_user, seed, step, rank, dt, next_dt) -> dict:
        """bh_step_resident: one whole time step, physics included, on the
        device.  Returns the report (energies, counts, stats)."""
        rep = ResidentReport()
        self._check(lib().bh_step_resident(self._h, n_user, seed, step, rank, dt, next_dt,
                                           C.byref(rep)))
        self._n_new = 0
        return rep.as_dict()

    def physics(self):
        """bh_get_physics: the opacities, Fleck factors and source energies
        of the last resident step (or of the last bh_set_mesh / energies)."""
        n, G = self.n_cells, self.n_groups
        sa, ss = np.zeros((n, G)), np.zeros((n, G))
        f, em, src = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(lib().bh_get_physics(self._h, _p(sa), _p(ss), _p(f), _p(em), _p(src)))
        return dict(sigma_a=sa, sigma_s=ss, fleck=f, emission_E=em, source_E=src)

    def step_resident_instances(self, n_user, step, rank, dt, next_dt):
        """bh_step_resident_instances: one step of every instance, physics
        included, on the device.  Returns (total report, per-instance
        structured array)."""
        k = len(self._inst_keep[1]) if getattr(self, "_inst_keep", None) else 0
        nu = np.ascontiguousarray(n_user, dtype=np.uint64)
        if len(nu) != k:
            raise ValueError("one photon count per instance")
        rep = ResidentReport()
        per = (InstanceReport * k)()
        self._check(lib().bh_step_resident_instances(self._h, _p(nu), step, rank, dt, next_dt,
                                                     C.byref(rep), per))
        self._n_new = 0
        names = [n for n, _ in InstanceReport._fields_]
        arr = np.zeros(k, dtype=[(n, np.uint64 if n == "n_new" else np.float64)
                                 for n in names])
        for n in names:
            arr[n] = [getattr(r, n) for r in per]
        return rep.as_dict(), arr

    def temperature(self):
        T_e, T_r = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(lib().bh_get_temperature(self._h, _p(T_e), _p(T_r)))
        return T_e, T_r

    # -- RNG probes ----------------------------------------------------------
    def rng_draws(self, seed, stream, n, ctr0=0):
        out = np.zeros(n)
        self._check(lib().bh_rng_draws(self._h, seed, stream, ctr0, n, _p(out)))
        return out

    def copy_probe(self, n_bytes: int, reps: int = 10) -> float:
        """Streaming copy rate in GB/s (read + write)."""
        g = C.c_double()
        self._check(lib().bh_copy_probe(self._h, n_bytes, reps, C.byref(g)))
        return g.value

    def threefry(self, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=np.uint64).reshape(-1, 4)
        out = np.zeros((len(w), 2), dtype=np.uint64)
        self._check(lib().bh_threefry(self._h, _p(w), len(w), _p(out)))
        return out


# ---------------------------------------------------------------------------
# Host driver (include/branson_host.h)
# ---------------------------------------------------------------------------
HOST_EXPORTS = [
    "bh_driver_create", "bh_driver_destroy", "bh_driver_last_error", "bh_driver_finished",
    "bh_driver_n_cells", "bh_driver_n_photons", "bh_driver_seed", "bh_driver_calculate_physics",
    "bh_driver_prepare", "bh_driver_transport", "bh_driver_finish", "bh_driver_step",
    "bh_driver_get_tallies", "bh_driver_get_temperature", "bh_driver_get_mesh",
    "bh_driver_get_physics", "bh_driver_engine",
]


class DriverConfig(C.Structure):
    _fields_ = [("engine", Options), ("device", C.c_int32), ("n_groups", C.c_uint32),
                ("rank", C.c_uint32), ("n_ranks", C.c_uint32), ("n_photons", C.c_uint64),
                ("t_stop", C.c_double), ("device_physics", C.c_int32), ("reserved", C.c_int32)]


class DriverReport(C.Structure):
    _fields_ = [("step", C.c_uint32), ("time", C.c_double), ("dt", C.c_double),
                ("next_dt", C.c_double), ("transported", C.c_uint64),
                ("census_size", C.c_uint64), ("n_segments", C.c_uint64),
                ("n_flushes", C.c_uint64), ("n_exit", C.c_uint64), ("n_killed", C.c_uint64),
                ("emission_E", C.c_double), ("source_E", C.c_double),
                ("absorbed_E", C.c_double), ("exit_E", C.c_double),
                ("pre_census_E", C.c_double), ("post_census_E", C.c_double),
                ("pre_mat_E", C.c_double), ("post_mat_E", C.c_double),
                ("rad_conservation", C.c_double), ("mat_conservation", C.c_double),
                ("source_ms", C.c_double), ("transport_ms", C.c_double),
                ("census_ms", C.c_double), ("prepare_ms", C.c_double),
                ("device_phase_ms", C.c_double), ("finish_ms", C.c_double),
                ("n_replayed", C.c_uint64), ("lattice", C.c_int32), ("tally", C.c_int32),
                ("replay_ms", C.c_double), ("n_allocs", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _host_lib():
    L = lib()
    if getattr(L, "_host_ready", False):
        return L
    P, U32, I = C.c_void_p, C.c_uint32, C.c_int
    L.bh_driver_create.argtypes = [C.c_char_p, P, P]
    L.bh_driver_destroy.argtypes = [P]
    L.bh_driver_destroy.restype = None
    L.bh_driver_last_error.argtypes = [P]
    L.bh_driver_last_error.restype = C.c_char_p
    for n in ("bh_driver_finished", "bh_driver_calculate_physics", "bh_driver_prepare",
              "bh_driver_transport"):
        getattr(L, n).argtypes = [P]
        getattr(L, n).restype = I
    L.bh_driver_n_cells.argtypes = [P]
    L.bh_driver_n_cells.restype = U32
    L.bh_driver_n_photons.argtypes = [P]
    L.bh_driver_n_photons.restype = C.c_uint64
    L.bh_driver_seed.argtypes = [P]
    L.bh_driver_seed.restype = U32
    L.bh_driver_finish.argtypes = [P, P]
    L.bh_driver_step.argtypes = [P, P]
    L.bh_driver_get_tallies.argtypes = [P, P, P]
    L.bh_driver_get_temperature.argtypes = [P, P, P]
    L.bh_driver_get_mesh.argtypes = [P, P, P, P, P]
    L.bh_driver_get_physics.argtypes = [P, P, P, P, P, P, P, P, P]
    L.bh_driver_engine.argtypes = [P]
    L.bh_driver_engine.restype = P
    for n in ("bh_driver_finish", "bh_driver_step", "bh_driver_get_tallies",
              "bh_driver_get_temperature", "bh_driver_get_mesh", "bh_driver_get_physics"):
        getattr(L, n).restype = I
    L._host_ready = True
    return L


class Driver:
    """Branson's replicated time loop over the engine (branson_host.h)."""

    def __init__(self, xml_path: str, device: int = 0, n_groups: int = 1, rank: int = 0,
                 n_ranks: int = 1, n_photons: int = 0, t_stop: float = 0.0,
                 precision: int = PRECISION_FP64, algorithm: int = ALGO_HISTORY,
                 tally_mode: int = 0, blocks_per_cu: int = 0, refill_threshold: int = 0,
                 reserve_photons: int = 0, fp32_tol: float = 0.0, block_size: int = 0,
                 batch_photons: int = 0, step_graph: bool = False,
                 general_mesh: bool = False, device_physics: bool = False,
                 source_mode: int = 0):
        L = _host_lib()
        self._d = C.c_void_p()
        cfg = DriverConfig(Options(precision, algorithm, block_size, blocks_per_cu, tally_mode,
                                   refill_threshold, reserve_photons, fp32_tol, batch_photons,
                                   int(step_graph), int(general_mesh), source_mode), device,
                           n_groups,
                           rank, n_ranks, n_photons, t_stop, int(device_physics), 0)
        rc = L.bh_driver_create(os.fsencode(xml_path), C.byref(cfg), C.byref(self._d))
        if rc != BH_OK:
            msg = L.bh_driver_last_error(self._d).decode()
            L.bh_driver_destroy(self._d)
            self._d = None
            raise BhError(rc, msg)
        self.n_cells = L.bh_driver_n_cells(self._d)
        self.n_groups = n_groups

    def close(self):
        if getattr(self, "_d", None):
            _host_lib().bh_driver_destroy(self._d)
            self._d = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != BH_OK:
            raise BhError(rc, _host_lib().bh_driver_last_error(self._d).decode())

    @property
    def n_photons(self):
        return _host_lib().bh_driver_n_photons(self._d)

    @property
    def seed(self):
        return _host_lib().bh_driver_seed(self._d)

    def finished(self) -> bool:
        return bool(_host_lib().bh_driver_finished(self._d))

    def calculate_physics(self):
        self._check(_host_lib().bh_driver_calculate_physics(self._d))

    def prepare(self):
        self._check(_host_lib().bh_driver_prepare(self._d))

    def transport(self):
        self._check(_host_lib().bh_driver_transport(self._d))

    def finish(self) -> dict:
        r = DriverReport()
        self._check(_host_lib().bh_driver_finish(self._d, C.byref(r)))
        return r.as_dict()

    def step(self) -> dict:
        r = DriverReport()
        self._check(_host_lib().bh_driver_step(self._d, C.byref(r)))
        return r.as_dict()

    def run(self) -> list:
        out = []
        while not self.finished():
            out.append(self.step())
        return out

    def tallies(self):
        a, t = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(_host_lib().bh_driver_get_tallies(self._d, _p(a), _p(t)))
        return a, t

    def temperature(self):
        te, tr = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(_host_lib().bh_driver_get_temperature(self._d, _p(te), _p(tr)))
        return te, tr

    def mesh(self):
        n = self.n_cells
        nodes = np.zeros((n, 6))
        bc = np.zeros((n, 6), dtype=np.int32)
        nxt = np.zeros((n, 6), dtype=np.uint32)
        reg = np.zeros(n, dtype=np.uint32)
        self._check(_host_lib().bh_driver_get_mesh(self._d, _p(nodes), _p(bc), _p(nxt),
                                                   _p(reg)))
        return nodes, bc, nxt, reg

    def physics(self):
        n, G = self.n_cells, self.n_groups
        sa, ss = np.zeros((n, G)), np.zeros((n, G))
        f, em, src, cen = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        face = np.zeros(n, dtype=np.int8)
        tot = np.zeros(1)
        self._check(_host_lib().bh_driver_get_physics(self._d, _p(sa), _p(ss), _p(f), _p(em),
                                                      _p(src), _p(cen), _p(face), _p(tot)))
        return dict(sigma_a=sa, sigma_s=ss, fleck=f, emission_E=em, source_E=src,
                    census_E=cen, source_face=face, total_E=float(tot[0]))
