"""Python binding of libmcpm.so, the MI355X slab particle step.

Mirrors the reference's per-step operator interface (src/cpp_headers.h):

    ============================  ==========================================
    reference (C++)               here
    ============================  ==========================================
    cell_update                   Simulation.cell_update()
    Collision_ESMC_new3           Simulation.collide_esmc(step)
    poisson_solver                Simulation.poisson_solver() (+ kick)
    stream_particles + apply_BC   Simulation.stream_bc(step)
    cell_poly_velocity            Simulation.sample()
    increase_collision_rate       contact_y(params, n)
    main.cpp:250-269 omega init   Simulation.init_omega_max(n_gas)
    ============================  ==========================================

The library is required: importing this module on a machine where
libmcpm.so is missing or no GPU is visible raises, there is no CPU path.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcpm.so")

FP64, FP32 = 0, 1
BC_DIFFUSE, BC_PERIODIC, BC_SPECULAR, BC_REINJECT = 0, 1, 2, 3
Y_REFERENCE, Y_CARNAHAN_STARLING = 0, 1
LR_NONE, LR_DIRECT, LR_DENSITY_EXPANSION, LR_SCREENED_POISSON = 0, 1, 2, 3
MODEL_ESMC, MODEL_FP_DENSE, MODEL_FP_IDEAL, MODEL_DSMC_VHS, MODEL_NONE = 0, 1, 2, 3, 4
MODEL_DSMC_HS, MODEL_CBA_VHS, MODEL_CBA_HS, MODEL_FP_LINEAR = 5, 6, 7, 8
MODEL_ESMC_NEW4 = 9
NMOM, NSAMP, NCLOS = 34, 26, 28

STATUS = {
    0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "capacity exceeded",
    4: "particle outside the domain", 5: "non-finite velocity",
    6: "too many boundary interactions in one step", 7: "call out of order",
}


class McpmError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {STATUS.get(code, 'unknown')} (status {code})")
        self.code = code


class CParams(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("y_model", ctypes.c_int32), ("long_range", ctypes.c_int32),
        ("n_cells", ctypes.c_int32), ("model", ctypes.c_int32),
        ("capacity", ctypes.c_int64), ("seed", ctypes.c_uint64),
        ("length", ctypes.c_double), ("wall_lo", ctypes.c_double), ("wall_hi", ctypes.c_double),
        ("T_wall_lo", ctypes.c_double), ("T_wall_hi", ctypes.c_double),
        ("U_wall_lo", ctypes.c_double), ("U_wall_hi", ctypes.c_double),
        ("sigma", ctypes.c_double), ("mass", ctypes.c_double), ("kb", ctypes.c_double),
        ("b", ctypes.c_double), ("Fn", ctypes.c_double), ("cell_volume", ctypes.c_double),
        ("n_ratio", ctypes.c_double), ("delta_t", ctypes.c_double),
        ("ast", ctypes.c_double), ("bst", ctypes.c_double), ("eqdist", ctypes.c_double),
        ("mu", ctypes.c_double), ("T0", ctypes.c_double), ("visp", ctypes.c_double),
        ("crref", ctypes.c_double), ("phi", ctypes.c_double),
        ("dims", ctypes.c_int32), ("n_cells_x", ctypes.c_int32),
        ("boundary_x", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("length_x", ctypes.c_double),
    ]


class CAccessRates(ctypes.Structure):
    _fields_ = [("gather_record", ctypes.c_double), ("gather_word", ctypes.c_double),
                ("atomic_min_word", ctypes.c_double), ("scatter_record", ctypes.c_double),
                ("scatter_word", ctypes.c_double), ("gather_byte", ctypes.c_double)]


class CAddRemoveStats(ctypes.Structure):
    _fields_ = [("before", ctypes.c_int64), ("split", ctypes.c_int64),
                ("created", ctypes.c_int64), ("deleted", ctypes.c_int64),
                ("after", ctypes.c_int64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class CFieldStats(ctypes.Structure):
    _fields_ = [("residual", ctypes.c_double), ("cg_iterations", ctypes.c_int32),
                ("used_network", ctypes.c_int32)]

    def as_dict(self):
        return {"residual": self.residual, "cg_iterations": self.cg_iterations,
                "used_network": bool(self.used_network)}


class CStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int64), ("valid", ctypes.c_int64),
                ("accepted", ctypes.c_int64), ("rounds", ctypes.c_int32),
                ("revisits", ctypes.c_int32)]

    def as_dict(self):
        return {"candidates": self.candidates, "valid": self.valid,
                "accepted": self.accepted, "rounds": self.rounds, "revisits": self.revisits}


@dataclass
class Params:
    """Slab (dims = 1) or plane (dims = 2) parameters; names follow the
    reference's GAS/BOX fields.  On the plane n_cells, length and boundary
    describe y and n_cells_x, length_x, boundary_x describe x."""
    n_cells: int
    length: float
    capacity: int
    precision: int = FP64
    boundary: int = BC_DIFFUSE
    y_model: int = Y_REFERENCE
    long_range: int = LR_NONE
    model: int = MODEL_ESMC
    seed: int = 1
    wall_lo: float = float("nan")
    wall_hi: float = float("nan")
    T_wall_lo: float = 1.0
    T_wall_hi: float = 1.0
    U_wall_lo: float = 0.0
    U_wall_hi: float = 0.0
    sigma: float = 1.0
    mass: float = 1.0
    kb: float = 1.0
    b: float = float("nan")
    Fn: float = 1.0
    cell_volume: float = float("nan")
    n_ratio: float = 1.0
    delta_t: float = 0.01
    ast: float = 0.0
    bst: float = 1.0
    eqdist: float = 1.0
    mu: float = float("nan")
    T0: float = 1.0
    visp: float = 0.5
    crref: float = 1.0
    phi: float = -4.0
    dims: int = 1
    n_cells_x: int = 1
    boundary_x: int = BC_SPECULAR
    reserved1: int = 0
    length_x: float = 1.0

    def __post_init__(self):
        if math.isnan(self.b):
            self.b = 2.0 * math.pi * self.sigma ** 3 / 3.0          # myIO.cpp:2828
        if math.isnan(self.wall_lo):
            self.wall_lo = self.sigma + 1e-15                       # Main_Functions.cpp:5126
        if math.isnan(self.wall_hi):
            self.wall_hi = self.length - self.sigma - 1e-15         # Main_Functions.cpp:5088
        if math.isnan(self.cell_volume):
            if self.dims == 2:
                # dx dy dz with unit depth: the injection problem scales Len[0]
                # and Len[1] by sigma but keeps Len[2] = 1 (myIO.cpp:2741-2761)
                self.cell_volume = (self.length_x / self.n_cells_x) * (self.length / self.n_cells) * 1.0
            else:
                self.cell_volume = self.length / self.n_cells
        if math.isnan(self.mu):
            # hard-sphere viscosity 5/(16 sigma^2) sqrt(m kB T0 / pi) (myIO.cpp:2829)
            self.mu = 5.0 / (16.0 * self.sigma ** 2) * math.sqrt(self.mass * self.kb * self.T0 / math.pi)

    @property
    def dy(self) -> float:
        return self.length / self.n_cells

    @property
    def dx(self) -> float:
        return self.length_x / self.n_cells_x

    @property
    def total_cells(self) -> int:
        return self.n_cells * (self.n_cells_x if self.dims == 2 else 1)

    @property
    def n_nodes(self) -> int:
        return (self.n_cells_x + 1) * (self.n_cells + 1)

    def to_c(self) -> CParams:
        c = CParams()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c

    @classmethod
    def from_c(cls, c: CParams) -> "Params":
        return cls(**{f.name: getattr(c, f.name) for f in fields(cls)})


class CReplica(ctypes.Structure):
    """mcpm_replica: one slab of a batched context."""
    _fields_ = [("n_cells", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("capacity", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("cell_volume", ctypes.c_double), ("Fn", ctypes.c_double)]


class CSettings(ctypes.Structure):
    """mcpm_settings: what readSettingsFile derives (myIO.cpp:1756-3016)."""
    _fields_ = [
        ("params", CParams), ("n_particles", ctypes.c_int64), ("n_drop", ctypes.c_int64),
        ("problem", ctypes.c_int32), ("num_steps", ctypes.c_int32), ("every", ctypes.c_int32),
        ("after", ctypes.c_int32), ("reset_every", ctypes.c_int32),
        ("num_realizations", ctypes.c_int32), ("init_step", ctypes.c_int32),
        ("direction", ctypes.c_int32 * 3), ("num_cells", ctypes.c_int32 * 3),
        ("reserved2", ctypes.c_int32), ("box_len", ctypes.c_double * 3),
        ("n", ctypes.c_double), ("T", ctypes.c_double), ("U0", ctypes.c_double * 3),
        ("T_init", ctypes.c_double), ("drop_dia", ctypes.c_double), ("epsilon", ctypes.c_double),
        ("lambda_", ctypes.c_double), ("Kn", ctypes.c_double), ("model", ctypes.c_char * 32),
        ("output", ctypes.c_char * 128),
    ]


def read_settings(path: str) -> CSettings:
    """mcpm_read_settings: a reference settings file -> CSettings (params
    ready for Simulation after Params.from_c)."""
    lib = load_library()
    s = CSettings()
    st = lib.mcpm_read_settings(os.fsencode(path), ctypes.byref(s))
    if st != 0:
        raise McpmError(st, "mcpm_read_settings")
    return s


def _u8p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


_lib = None


def load_library() -> ctypes.CDLL:
    """Load libmcpm.so; raise if it is missing (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmcpm.so not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dpp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
    u32pp = ctypes.POINTER(ctypes.c_uint32)
    fpp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "mcpm_create": (i32, [ctypes.POINTER(CParams), ctypes.POINTER(vp)]),
        "mcpm_destroy": (i32, [vp]),
        "mcpm_sync": (i32, [vp]),
        "mcpm_upload": (i32, [vp, i64, dpp, dpp, dpp, dpp]),
        "mcpm_download": (i32, [vp, dpp, dpp, dpp, dpp]),
        "mcpm_upload_xy": (i32, [vp, i64, dpp, dpp, dpp, dpp, dpp]),
        "mcpm_download_xy": (i32, [vp, dpp, dpp, dpp, dpp, dpp]),
        "mcpm_get_nodes": (i32, [vp, dpp, dpp]),
        "mcpm_particle_count": (i64, [vp]),
        "mcpm_cell_update": (i32, [vp]),
        "mcpm_collide_esmc": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_collide_esmc_new4": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_track_ids": (i32, [vp, i32]),
        "mcpm_download_ids": (i32, [vp, ctypes.POINTER(i32)]),
        "mcpm_set_cell_ratio": (i32, [vp, ctypes.POINTER(i32)]),
        "mcpm_set_particle_ratio": (i32, [vp, ctypes.POINTER(i32)]),
        "mcpm_get_particle_ratio": (i32, [vp, ctypes.POINTER(i32)]),
        "mcpm_addremove": (i32, [vp, i32, i64, ctypes.POINTER(CAddRemoveStats)]),
        "mcpm_poisson_solver": (i32, [vp]),
        "mcpm_kick": (i32, [vp]),
        "mcpm_stream_bc": (i32, [vp, i64, i32]),
        "mcpm_step": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_graph_capture": (i32, [vp, i64, i32]),
        "mcpm_graph_steps": (i32, [vp, i64, i64, ctypes.POINTER(CStats)]),
        "mcpm_graph_release": (i32, [vp]),
        "mcpm_init_omega_max": (i32, [vp, ctypes.c_double, i32]),
        "mcpm_set_omega_max": (i32, [vp, dpp]),
        "mcpm_get_cells": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), dpp, dpp, dpp]),
        "mcpm_moments": (i32, [vp, dpp]),
        "mcpm_sample": (i32, [vp]),
        "mcpm_get_samples": (i32, [vp, dpp]),
        "mcpm_reset_samples": (i32, [vp]),
        "mcpm_measure_pressure": (i32, [vp, i32]),
        "mcpm_get_escaped": (i32, [vp, ctypes.POINTER(i64)]),
        "mcpm_energy_momentum": (i32, [vp, dpp]),
        "mcpm_energy_momentum_noalloc": (i32, [vp, vp, vp, i64, ctypes.c_double, vp, vp]),
        "mcpm_thermostat": (i32, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double]),
        "mcpm_read_settings": (i32, [ctypes.c_char_p, ctypes.POINTER(CSettings)]),
        "mcpm_create_batch": (i32, [ctypes.POINTER(CParams), i32, ctypes.POINTER(CReplica),
                                    ctypes.POINTER(vp)]),
        "mcpm_upload_batch": (i32, [vp, ctypes.POINTER(i64), dpp, dpp, dpp, dpp]),
        "mcpm_init_omega_max_batch": (i32, [vp, dpp, i32]),
        "mcpm_initialize": (i32, [vp, ctypes.POINTER(CSettings)]),
        "mcpm_write_postproc": (i32, [vp, ctypes.POINTER(CSettings), i64, ctypes.c_char_p]),
        "mcpm_get_pressure": (i32, [vp, dpp, dpp, dpp]),
        "mcpm_contact_y": (ctypes.c_double, [ctypes.POINTER(CParams), ctypes.c_double]),
        "mcpm_debug_collision_log": (i32, [vp, i32]),
        "mcpm_get_collision_log": (i32, [vp, i64, u32pp, u32pp, u32pp,
                                         ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(i64)]),
        "mcpm_timing_enable": (i32, [vp, i32]),
        "mcpm_timing_get": (i32, [vp, i32, dpp, ctypes.POINTER(i64)]),
        "mcpm_status_string": (ctypes.c_char_p, [i32]),
        "mcpm_fp_closures": (i32, [vp, i64]),
        "mcpm_velocity_fp": (i32, [vp, i64]),
        "mcpm_get_closures": (i32, [vp, dpp]),
        "mcpm_collide_dsmc_vhs": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_collide_dsmc_hs": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_collide_cba_vhs": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_collide_cba_hs": (i32, [vp, i64, ctypes.POINTER(CStats)]),
        "mcpm_vlasov_integral": (i32, [vp]),
        "mcpm_density_expansion": (i32, [vp]),
        "mcpm_get_kde": (i32, [vp, dpp, dpp, dpp]),
        "mcpm_set_tuning": (i32, [vp, i32, i64]),
        "mcpm_get_tuning": (i32, [vp, i32, ctypes.POINTER(i64)]),
        "mcpm_probe_random_access": (i32, [i64, i32, ctypes.POINTER(CAccessRates)]),
        "mcpm_fno_param_count": (i32, [i32, ctypes.POINTER(i64)]),
        "mcpm_fno_create": (i32, [i32, i32, i32, fpp, ctypes.POINTER(vp)]),
        "mcpm_fno_destroy": (i32, [vp]),
        "mcpm_fno_set_params": (i32, [vp, fpp]),
        "mcpm_fno_get_params": (i32, [vp, fpp]),
        "mcpm_fno_set_norm": (i32, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double]),
        "mcpm_fno_forward": (i32, [vp, fpp, fpp]),
        "mcpm_fno_loss_grad": (i32, [vp, fpp, fpp, ctypes.c_float, fpp, fpp, fpp]),
        "mcpm_fno_reserve_snapshots": (i32, [vp, i32]),
        "mcpm_fno_add_snapshot_host": (i32, [vp, fpp, fpp, ctypes.c_float]),
        "mcpm_fno_add_snapshot": (i32, [vp, vp, ctypes.c_float]),
        "mcpm_fno_snapshot_count": (i32, [vp, ctypes.POINTER(i32)]),
        "mcpm_fno_train": (i32, [vp, i32, ctypes.c_float, ctypes.c_uint64, i32, fpp]),
        "mcpm_fno_field": (i32, [vp, vp, ctypes.c_double, i32, ctypes.POINTER(CFieldStats)]),
        "mcpm_set_field_surrogate": (i32, [vp, vp, ctypes.c_double, i64]),
        "mcpm_get_field_surrogate_stats": (i32, [vp, ctypes.POINTER(CFieldStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def contact_y(p: Params, n: float) -> float:
    """increase_collision_rate (Main_Functions.cpp:1906-1915)."""
    lib = load_library()
    cp = p.to_c()
    return lib.mcpm_contact_y(ctypes.byref(cp), float(n))


def energy_momentum_noalloc(u1_ptr: int, u2_ptr: int, u3_ptr: int, n: int, mass_fn: float,
                            out_ptr: int, stream: int = 0) -> None:
    """mcpm_energy_momentum_noalloc on device addresses (u1, u2, u3: n FP64
    velocities; out: 4 doubles); asynchronous on `stream` (0: default)."""
    lib = load_library()
    st = lib.mcpm_energy_momentum_noalloc(u1_ptr, u2_ptr, u3_ptr, int(n), float(mass_fn), out_ptr,
                                          stream or None)
    if st:
        raise McpmError(st, "mcpm_energy_momentum_noalloc")


def probe_random_access(n_records: int, record_bytes: int) -> dict:
    """Random-access rates of the current device (accesses per second) over
    n_records records of record_bytes, n_records 4-byte words and n_records
    bytes: record gathers, word gathers, 32-bit atomic minimums, record
    stores, word stores, byte gathers."""
    lib = load_library()
    r = CAccessRates()
    st = lib.mcpm_probe_random_access(int(n_records), int(record_bytes), ctypes.byref(r))
    if st != 0:
        raise McpmError(st, "probe_random_access")
    return {f: getattr(r, f) for f, _ in CAccessRates._fields_}


TIMER_COLLIDE, TIMER_SORT, TIMER_GEN, TIMER_FIELD, TIMER_STEP = 0, 1, 2, 3, 4
TIMER_FP_VELOCITY, TIMER_FP_CLOSURES = 5, 6
TUNE_BANDS, TUNE_ROUND_GRID, TUNE_SORT_CHUNK = 0, 1, 2
TUNE_SORT_WINDOW, TUNE_SORT_WINDOW_RUNS, TUNE_SORT_WINDOW_REDONE = 3, 4, 5
TUNE_ROUND_UNROLL = 6


class Simulation:
    """A device-resident slab of particles driven step by step."""

    def __init__(self, params: Params):
        self.params = params
        self._lib = load_library()
        self._cp = params.to_c()
        h = ctypes.c_void_p()
        self._check(self._lib.mcpm_create(ctypes.byref(self._cp), ctypes.byref(h)), "mcpm_create")
        self._h = h

    @classmethod
    def batch(cls, common: Params, replicas) -> "Simulation":
        """mcpm_create_batch: independent slabs in one context; replicas is a
        list of dicts with n_cells, capacity, seed, cell_volume, Fn."""
        self = cls.__new__(cls)
        self._lib = load_library()
        reps = (CReplica * len(replicas))()
        total = 0
        for r, d in enumerate(replicas):
            reps[r].n_cells = d["n_cells"]
            reps[r].capacity = d["capacity"]
            reps[r].seed = d["seed"]
            reps[r].cell_volume = d["cell_volume"]
            reps[r].Fn = d["Fn"]
            total += d["n_cells"]
        self._reps = reps
        self.params = Params(**{f.name: getattr(common, f.name) for f in fields(Params)})
        self.params.n_cells = total
        self.params.capacity = sum(d["capacity"] for d in replicas)
        self._cp = common.to_c()
        h = ctypes.c_void_p()
        self._check(self._lib.mcpm_create_batch(ctypes.byref(self._cp), len(replicas), reps,
                                                ctypes.byref(h)), "mcpm_create_batch")
        self._h = h
        return self

    def upload_batch(self, counts, y, u1, u2, u3):
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (y, u1, u2, u3)]
        self._check(self._lib.mcpm_upload_batch(
            self._h, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            *[_dp(a) for a in arrs]), "mcpm_upload_batch")

    def init_omega_max_batch(self, n_gas, times_y: bool = True):
        n_gas = np.ascontiguousarray(n_gas, dtype=np.float64)
        self._check(self._lib.mcpm_init_omega_max_batch(self._h, _dp(n_gas), 1 if times_y else 0),
                    "mcpm_init_omega_max_batch")

    # -- plumbing -----------------------------------------------------------
    def _check(self, st: int, what: str):
        if st != 0:
            raise McpmError(st, what)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mcpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_cells(self) -> int:
        """Total number of cells (N0 N1 on the plane)."""
        return self.params.total_cells

    # -- data ---------------------------------------------------------------
    def upload(self, y, u1, u2, u3):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (y, u1, u2, u3)]
        n = arrs[0].shape[0]
        self._check(self._lib.mcpm_upload(self._h, n, *[_dp(a) for a in arrs]), "mcpm_upload")

    def download(self):
        n = self._lib.mcpm_particle_count(self._h)
        out = [np.empty(n, np.float64) for _ in range(4)]
        self._check(self._lib.mcpm_download(self._h, *[_dp(a) for a in out]), "mcpm_download")
        return tuple(out)

    def upload_xy(self, x, y, u1, u2, u3):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u1, u2, u3)]
        n = arrs[0].shape[0]
        self._check(self._lib.mcpm_upload_xy(self._h, n, *[_dp(a) for a in arrs]), "mcpm_upload_xy")

    def download_xy(self):
        n = self._lib.mcpm_particle_count(self._h)
        out = [np.empty(n, np.float64) for _ in range(5)]
        self._check(self._lib.mcpm_download_xy(self._h, *[_dp(a) for a in out]), "mcpm_download_xy")
        return tuple(out)

    def nodes(self):
        """(n_node, potential) on the plane's (N0+1)(N1+1) nodes, shape (N1+1, N0+1)."""
        nn = self.params.n_nodes
        nd = np.empty(nn)
        pot = np.empty(nn)
        self._check(self._lib.mcpm_get_nodes(self._h, _dp(nd), _dp(pot)), "mcpm_get_nodes")
        shape = (self.params.n_cells + 1, self.params.n_cells_x + 1)
        return nd.reshape(shape), pot.reshape(shape)

    def cells(self):
        nc = self.n_cells
        cnt = np.empty(nc, np.int32)
        st = np.empty(nc, np.int32)
        n = np.empty(nc)
        om = np.empty(nc)
        f2 = np.empty(nc)
        self._check(self._lib.mcpm_get_cells(self._h, _i32p(cnt), _i32p(st), _dp(n), _dp(om),
                                             _dp(f2)), "mcpm_get_cells")
        return {"count": cnt, "start": st, "n": n, "omega_max": om, "F2": f2}

    # -- operators ----------------------------------------------------------
    def cell_update(self):
        self._check(self._lib.mcpm_cell_update(self._h), "mcpm_cell_update")

    def collide_esmc(self, step: int) -> dict:
        s = CStats()
        self._check(self._lib.mcpm_collide_esmc(self._h, step, ctypes.byref(s)), "mcpm_collide_esmc")
        return s.as_dict()

    def track_ids(self, enable: bool = True):
        self._check(self._lib.mcpm_track_ids(self._h, 1 if enable else 0), "mcpm_track_ids")

    def ids(self) -> np.ndarray:
        """Upload slot of the particle at each position of download()."""
        out = np.empty(int(self._lib.mcpm_particle_count(self._h)), dtype=np.int32)
        self._check(self._lib.mcpm_download_ids(self._h, _i32p(out)), "mcpm_download_ids")
        return out

    def set_cell_ratio(self, ratio):
        """Per-cell particle weights cells[num].n_ratio (None: p.n_ratio
        everywhere); every particle takes its cell's weight."""
        if ratio is None:
            self._check(self._lib.mcpm_set_cell_ratio(self._h, None), "mcpm_set_cell_ratio")
            return
        r = np.ascontiguousarray(ratio, dtype=np.int32)
        self._check(self._lib.mcpm_set_cell_ratio(self._h, _i32p(r)), "mcpm_set_cell_ratio")

    def set_particle_ratio(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        self._check(self._lib.mcpm_set_particle_ratio(self._h, _i32p(w)),
                    "mcpm_set_particle_ratio")

    def particle_ratio(self) -> np.ndarray:
        out = np.empty(int(self._lib.mcpm_particle_count(self._h)), dtype=np.int32)
        self._check(self._lib.mcpm_get_particle_ratio(self._h, _i32p(out)),
                    "mcpm_get_particle_ratio")
        return out

    def addremove(self, n_ratio: int, step: int) -> dict:
        """inverted_addremove_particles with gas n_ratio = n_ratio."""
        s = CAddRemoveStats()
        self._check(self._lib.mcpm_addremove(self._h, int(n_ratio), int(step), ctypes.byref(s)),
                    "mcpm_addremove")
        return s.as_dict()

    def collide_esmc_new4(self, step: int) -> dict:
        s = CStats()
        self._check(self._lib.mcpm_collide_esmc_new4(self._h, step, ctypes.byref(s)),
                    "mcpm_collide_esmc_new4")
        return s.as_dict()

    def poisson_solver(self):
        self._check(self._lib.mcpm_poisson_solver(self._h), "mcpm_poisson_solver")

    def kick(self):
        self._check(self._lib.mcpm_kick(self._h), "mcpm_kick")

    def vlasov_integral(self):
        self._check(self._lib.mcpm_vlasov_integral(self._h), "mcpm_vlasov_integral")

    def density_expansion(self):
        self._check(self._lib.mcpm_density_expansion(self._h), "mcpm_density_expansion")

    def kde(self):
        nc = self.n_cells
        out = [np.empty(nc) for _ in range(3)]
        self._check(self._lib.mcpm_get_kde(self._h, *[_dp(a) for a in out]), "mcpm_get_kde")
        return tuple(out)

    def stream_bc(self, step: int, with_kick: bool = True):
        self._check(self._lib.mcpm_stream_bc(self._h, step, 1 if with_kick else 0), "mcpm_stream_bc")

    def step(self, step: int) -> dict:
        s = CStats()
        self._check(self._lib.mcpm_step(self._h, step, ctypes.byref(s)), "mcpm_step")
        return s.as_dict()

    def graph_capture(self, op_capacity: int = 0, rounds: int = 0):
        """Record one step as HIP graphs (mcpm_graph_capture)."""
        self._check(self._lib.mcpm_graph_capture(self._h, op_capacity, rounds),
                    "mcpm_graph_capture")

    def graph_steps(self, step0: int, nsteps: int) -> dict:
        """Steps step0 .. step0 + nsteps - 1 by graph replays; sums of the
        per-step statistics (mcpm_graph_steps)."""
        s = CStats()
        self._check(self._lib.mcpm_graph_steps(self._h, step0, nsteps, ctypes.byref(s)),
                    "mcpm_graph_steps")
        return s.as_dict()

    def graph_release(self):
        self._check(self._lib.mcpm_graph_release(self._h), "mcpm_graph_release")

    def sync(self):
        self._check(self._lib.mcpm_sync(self._h), "mcpm_sync")

    def fp_closures(self, step: int):
        self._check(self._lib.mcpm_fp_closures(self._h, step), "mcpm_fp_closures")

    def velocity_fp(self, step: int):
        self._check(self._lib.mcpm_velocity_fp(self._h, step), "mcpm_velocity_fp")

    def closures(self) -> np.ndarray:
        out = np.empty(self.n_cells * NCLOS)
        self._check(self._lib.mcpm_get_closures(self._h, _dp(out)), "mcpm_get_closures")
        return out.reshape(self.n_cells, NCLOS)

    def _collide(self, name: str, step: int) -> dict:
        s = CStats()
        self._check(getattr(self._lib, name)(self._h, step, ctypes.byref(s)), name)
        return s.as_dict()

    def collide_dsmc_vhs(self, step: int) -> dict:
        """Collision_DSMC_VHS (collisions.cpp:2063-2141)."""
        return self._collide("mcpm_collide_dsmc_vhs", step)

    def collide_dsmc_hs(self, step: int) -> dict:
        """Collision_DSMC_HS (collisions.cpp:1989-2061)."""
        return self._collide("mcpm_collide_dsmc_hs", step)

    def collide_cba_vhs(self, step: int) -> dict:
        """Collision_CBA_VHS (collisions.cpp:2144-2248)."""
        return self._collide("mcpm_collide_cba_vhs", step)

    def collide_cba_hs(self, step: int) -> dict:
        """Collision_CBA_HS (collisions.cpp:2251-2346)."""
        return self._collide("mcpm_collide_cba_hs", step)

    def init_omega_max(self, n_gas: float, times_y: bool = True):
        self._check(self._lib.mcpm_init_omega_max(self._h, float(n_gas), 1 if times_y else 0),
                    "mcpm_init_omega_max")

    def set_omega_max(self, om):
        om = np.ascontiguousarray(om, dtype=np.float64)
        if om.size != self.n_cells:
            raise ValueError(f"omega_max needs {self.n_cells} values, got {om.size}")
        self._check(self._lib.mcpm_set_omega_max(self._h, _dp(om)), "mcpm_set_omega_max")

    def moments(self) -> np.ndarray:
        out = np.empty(self.n_cells * NMOM)
        self._check(self._lib.mcpm_moments(self._h, _dp(out)), "mcpm_moments")
        return out.reshape(self.n_cells, NMOM)

    def compute_moments(self):
        """The cell moments on the device only (cell_update's every step)."""
        self._check(self._lib.mcpm_moments(self._h, None), "mcpm_moments")

    def sample(self):
        self._check(self._lib.mcpm_sample(self._h), "mcpm_sample")

    def samples(self) -> np.ndarray:
        out = np.empty(self.n_cells * NSAMP)
        self._check(self._lib.mcpm_get_samples(self._h, _dp(out)), "mcpm_get_samples")
        return out.reshape(self.n_cells, NSAMP)

    def initialize(self, settings: "CSettings"):
        """initialization (Main_Functions.cpp:4222-4815) on the device."""
        self._check(self._lib.mcpm_initialize(self._h, ctypes.byref(settings)), "mcpm_initialize")

    def write_postproc(self, settings: "CSettings", step: int, directory: str):
        """write_post_processing's PostProc file (myIO.cpp:597-897)."""
        self._check(self._lib.mcpm_write_postproc(self._h, ctypes.byref(settings), step,
                                                  os.fsencode(directory)),
                    "mcpm_write_postproc")

    def escaped(self) -> int:
        """Particles the evaporation boundary (BC_REINJECT) has re-inserted."""
        v = ctypes.c_int64()
        self._check(self._lib.mcpm_get_escaped(self._h, ctypes.byref(v)), "mcpm_get_escaped")
        return v.value

    def energy_momentum(self) -> np.ndarray:
        """FP64 kinetic energy (1/2) m Fn sum |U|^2 and momentum m Fn sum U."""
        out = np.empty(4)
        self._check(self._lib.mcpm_energy_momentum(self._h, _dp(out)), "mcpm_energy_momentum")
        return out

    def thermostat(self, zone_lo: float, T_lo: float, zone_hi: float, T_hi: float):
        """thermostat of the inverted problem (Main_Functions.cpp:1760-1830)."""
        self._check(self._lib.mcpm_thermostat(self._h, zone_lo, T_lo, zone_hi, T_hi),
                    "mcpm_thermostat")

    def measure_pressure(self, enable: bool = True):
        """measure_pressure (Main_Functions.cpp:13-34) on every moving push."""
        self._check(self._lib.mcpm_measure_pressure(self._h, 1 if enable else 0),
                    "mcpm_measure_pressure")

    def pressure(self) -> dict:
        """PostProc's p_m, p_rep and ptot columns (myIO.cpp:855, 889-890)."""
        out = {k: np.empty(self.n_cells) for k in ("p_m", "p_rep", "p_tot")}
        self._check(self._lib.mcpm_get_pressure(self._h, _dp(out["p_m"]), _dp(out["p_rep"]),
                                                _dp(out["p_tot"])), "mcpm_get_pressure")
        return out

    def reset_samples(self):
        self._check(self._lib.mcpm_reset_samples(self._h), "mcpm_reset_samples")

    # -- diagnostics --------------------------------------------------------
    def debug_collision_log(self, enable: bool = True):
        self._check(self._lib.mcpm_debug_collision_log(self._h, 1 if enable else 0), "debug_log")

    def collision_log(self, cap: int):
        pr = np.empty(cap, np.uint32)
        i = np.empty(cap, np.uint32)
        j = np.empty(cap, np.uint32)
        a = np.empty(cap, np.uint8)
        n = ctypes.c_int64()
        self._check(self._lib.mcpm_get_collision_log(self._h, cap, _u32p(pr), _u32p(i), _u32p(j),
                                                     _u8p(a), ctypes.byref(n)), "collision_log")
        k = n.value
        return pr[:k], i[:k], j[:k], a[:k]

    def set_tuning(self, knob: int, value: int):
        self._check(self._lib.mcpm_set_tuning(self._h, knob, int(value)), "set_tuning")

    def get_tuning(self, knob: int) -> int:
        v = ctypes.c_int64()
        self._check(self._lib.mcpm_get_tuning(self._h, knob, ctypes.byref(v)), "get_tuning")
        return v.value

    def timing_enable(self, on: bool = True):
        self._check(self._lib.mcpm_timing_enable(self._h, 1 if on else 0), "timing_enable")

    def timing(self, which: int):
        ms = ctypes.c_double()
        calls = ctypes.c_int64()
        self._check(self._lib.mcpm_timing_get(self._h, which, ctypes.byref(ms), ctypes.byref(calls)),
                    "timing_get")
        return ms.value, calls.value

    def set_field_surrogate(self, fno: "FieldSurrogate | None", accept_residual: float = 1e-3,
                            every: int = 50):
        """Attach the field surrogate to mcpm_step / poisson_solver (None detaches)."""
        h = fno._h if fno is not None else None
        self._check(self._lib.mcpm_set_field_surrogate(self._h, h, float(accept_residual),
                                                       int(every)), "set_field_surrogate")
        self._fno = fno

    def field_surrogate_stats(self) -> dict:
        st = CFieldStats()
        self._check(self._lib.mcpm_get_field_surrogate_stats(self._h, ctypes.byref(st)),
                    "get_field_surrogate_stats")
        return st.as_dict()


def plane_force(pot, h):
    """Gradient of the Q1 interpolant of a nodal field at the cell centres
    (FEM.cpp:570-583 evaluated at xi = eta = 1/2): shape (2, N1, N0)."""
    gx = 0.5 * ((pot[:-1, 1:] - pot[:-1, :-1]) + (pot[1:, 1:] - pot[1:, :-1])) / h
    gy = 0.5 * ((pot[1:, :-1] - pot[:-1, :-1]) + (pot[1:, 1:] - pot[:-1, 1:])) / h
    return np.stack([gx, gy])


def _periodic_nodes(a):
    """(N1, N0) periodic node values -> (N1 + 1, N0 + 1) with the seam rows."""
    out = np.empty((a.shape[0] + 1, a.shape[1] + 1))
    out[:-1, :-1] = a
    out[-1, :-1] = a[0]
    out[:-1, -1] = a[:, 0]
    out[-1, -1] = a[0, 0]
    return out


def plane_core(n, h, lam, sigma):
    """The core subtraction of compute_SP_2D on a periodic plane
    (FEM.cpp:497-554): core0 n + core2 (sum of the 4 neighbours - 4 n), the
    term the field loses, for (N1, N0) node densities."""
    ls = lam * sigma
    core0 = (1.0 - (ls + 1.0) * math.exp(-ls)) / lam ** 2
    core2 = ((-ls ** 3 / 6 - ls ** 2 / 2 - ls + math.exp(ls) - 1.0) * math.exp(-ls) / lam ** 4) / h ** 2
    nb = np.roll(n, 1, 0) + np.roll(n, -1, 0) + np.roll(n, 1, 1) + np.roll(n, -1, 1)
    return core0 * n + core2 * (nb - 4.0 * n)


def spectral_force_floor(n_node, h, lam, sigma, modes):
    """The smallest force error a spectral operator with `modes` retained
    modes per axis can reach on this density: exact on the retained modes
    (|kx|, |ky| < modes), the best pointwise gain on the others, the exact
    core term on both.  The field is the periodic Q1 solution of
    (K + lambda^2 M) phi = M n (FEM.cpp:73-110 element matrices, square cells
    of side h) minus the core term; the force the gradient of its bilinear
    interpolant at the cell centres (plane_force).  n_node: the
    (N1 + 1, N0 + 1) node densities of a periodic plane as Simulation.nodes
    returns them (the last row and column repeat the first).  Returns the
    relative force error of that best operator, of the retained modes alone,
    and the best gain."""
    n = np.asarray(n_node, dtype=np.float64)
    if n.ndim != 2 or min(n.shape) < 3:
        raise ValueError("node densities must be an (N1 + 1, N0 + 1) array")
    n = n[:-1, :-1]
    n1, n0 = n.shape
    cx = np.cos(2 * np.pi * np.fft.fftfreq(n0))[None, :]
    cy = np.cos(2 * np.pi * np.fft.fftfreq(n1))[:, None]
    K = 8.0 / 3.0 - (2.0 / 3.0) * (cx + cy) - (4.0 / 3.0) * cx * cy
    M = h * h * (2 + cx) * (2 + cy) / 9.0
    T = M / (K + lam * lam * M)
    nh = np.fft.fft2(n)
    core = plane_core(n, h, lam, sigma)

    def force(phi):
        return plane_force(_periodic_nodes(phi), h)

    exact = force(np.real(np.fft.ifft2(T * nh)) - core)
    den = np.linalg.norm(exact)
    kx = np.abs(np.fft.fftfreq(n0) * n0)[None, :]
    ky = np.abs(np.fft.fftfreq(n1) * n1)[:, None]
    keep = (kx < modes) & (ky < modes)
    low = force(np.real(np.fft.ifft2(np.where(keep, T * nh, 0.0))) - core)
    high = force(np.real(np.fft.ifft2(np.where(keep, 0.0, nh))))
    rest = exact - low
    hh = float(np.sum(high * high))
    gain = float(np.sum(high * rest) / hh) if hh > 0 else 0.0
    return {"floor": float(np.linalg.norm(rest - gain * high) / den),
            "modes_only": float(np.linalg.norm(rest) / den), "gain": gain}


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class FieldSurrogate:
    """Spectral neural operator standing in for the periodic plane's
    screened-Poisson solve (compute_SP_2D, FEM.cpp:249-610); no counterpart
    in the reference.  Width 64, four Fourier layers, `modes` retained modes
    per axis; trained on the GPU by the library's own forward, backward and
    Adam kernels on snapshots that the exact solver writes."""

    C = 64

    def __init__(self, n0: int, n1: int, modes: int, params: np.ndarray):
        self._lib = load_library()
        self.n0, self.n1, self.modes = n0, n1, modes
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != self.param_count(modes):
            raise ValueError(f"parameter vector of {params.size}, expected {self.param_count(modes)}")
        h = ctypes.c_void_p()
        st = self._lib.mcpm_fno_create(n0, n1, modes, _fp(params), ctypes.byref(h))
        if st:
            raise McpmError(st, "mcpm_fno_create")
        self._h = h

    @staticmethod
    def init_params(modes: int, seed: int = 0) -> np.ndarray:
        """Initial FP32 parameters (numpy PCG64 stream of `seed`): lift weight
        and bias U(-1, 1); per layer W and b U(-1/8, 1/8) (1/sqrt(64)) and the
        spectral weights U(-1, 1)/64^2 per component; projection weight
        U(-1/8, 1/8), bias 0.  Layout as in include/mcpm.h."""
        C = FieldSurrogate.C
        rng = np.random.default_rng(seed)
        rsz = 2 * modes * modes * C * C * 2
        out = [rng.uniform(-1, 1, C), rng.uniform(-1, 1, C)]
        a = 1.0 / np.sqrt(C)
        for _ in range(4):
            out += [rng.uniform(-a, a, C * C), rng.uniform(-a, a, C), rng.uniform(-1, 1, rsz) / (C * C)]
        out += [rng.uniform(-a, a, C), np.zeros(1)]
        v = np.concatenate(out).astype(np.float32)
        n = FieldSurrogate.param_count(modes)
        return np.concatenate([v, np.zeros(n - v.size, np.float32)])

    @staticmethod
    def param_count(modes: int) -> int:
        n = ctypes.c_int64()
        st = load_library().mcpm_fno_param_count(modes, ctypes.byref(n))
        if st:
            raise McpmError(st, "mcpm_fno_param_count")
        return n.value

    def _check(self, st: int, what: str):
        if st:
            raise McpmError(st, what)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mcpm_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def P(self) -> int:
        return self.n0 * self.n1

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._lib.mcpm_fno_set_params(self._h, _fp(params)), "fno_set_params")

    def params(self) -> np.ndarray:
        out = np.empty(self.param_count(self.modes), np.float32)
        self._check(self._lib.mcpm_fno_get_params(self._h, _fp(out)), "fno_get_params")
        return out

    def set_norm(self, in_shift, in_scale, out_shift, out_scale):
        self.norm = (in_shift, in_scale, out_shift, out_scale)
        self._check(self._lib.mcpm_fno_set_norm(self._h, float(in_shift), float(in_scale),
                                                float(out_shift), float(out_scale)), "fno_set_norm")

    def forward(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        y = np.empty(self.P, np.float32)
        self._check(self._lib.mcpm_fno_forward(self._h, _fp(u), _fp(y)), "fno_forward")
        return y.reshape(self.n1, self.n0)

    def loss_grad(self, u, t, w0: float):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        loss = ctypes.c_float()
        g = np.empty(self.param_count(self.modes), np.float32)
        y = np.empty(self.P, np.float32)
        self._check(self._lib.mcpm_fno_loss_grad(self._h, _fp(u), _fp(t), w0, ctypes.byref(loss),
                                                 _fp(g), _fp(y)), "fno_loss_grad")
        return loss.value, g, y.reshape(self.n1, self.n0)

    def reserve_snapshots(self, cap: int):
        self._check(self._lib.mcpm_fno_reserve_snapshots(self._h, cap), "fno_reserve_snapshots")

    def add_snapshot_host(self, u, t, w0: float):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        self._check(self._lib.mcpm_fno_add_snapshot_host(self._h, _fp(u), _fp(t), w0),
                    "fno_add_snapshot_host")

    def add_snapshot(self, sim: "Simulation", w0: float):
        self._check(self._lib.mcpm_fno_add_snapshot(self._h, sim._h, w0), "fno_add_snapshot")

    def snapshot_count(self) -> int:
        n = ctypes.c_int32()
        self._check(self._lib.mcpm_fno_snapshot_count(self._h, ctypes.byref(n)), "fno_snapshot_count")
        return n.value

    def train(self, iters: int, lr: float, seed: int = 0, n_snap: int = 0) -> np.ndarray:
        out = np.empty(max(iters, 1), np.float32)
        self._check(self._lib.mcpm_fno_train(self._h, iters, lr, seed, n_snap, _fp(out)),
                    "fno_train")
        return out[:iters]

    def field(self, sim: "Simulation", accept_residual: float = -1.0, exact: bool = False) -> dict:
        st = CFieldStats()
        self._check(self._lib.mcpm_fno_field(self._h, sim._h, float(accept_residual),
                                             1 if exact else 0, ctypes.byref(st)), "fno_field")
        return st.as_dict()
