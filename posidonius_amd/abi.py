"""ctypes mirror of include/posidonius_hip.h and the loader of the HIP library.

The structures here are field-for-field copies of the C declarations; the
constants carry the same names without the ``POS_`` prefix.  Loading the
library is explicit and loud: there is no CPU fallback for the integration
step.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ABI_VERSION = 4
MAX_BODIES = 16

OK = 0
E_INVALID = -1
E_HIP = -2
E_NO_DEVICE = -3
E_UNSUPPORTED = -4
E_LIMIT = -5

COORD_JACOBI = 0
COORD_DEMOCRATIC_HELIOCENTRIC = 1
COORD_WHDS = 2
COORDINATES = {"Jacobi": COORD_JACOBI, "DemocraticHeliocentric": COORD_DEMOCRATIC_HELIOCENTRIC, "WHDS": COORD_WHDS}

GR_KIDDER1995 = 0
GR_ANDERSON1975 = 1
GR_NEWHALL1983 = 2
GR_DISABLED = 3
GR_IMPLEMENTATIONS = {"Kidder1995": GR_KIDDER1995, "Anderson1975": GR_ANDERSON1975,
                      "Newhall1983": GR_NEWHALL1983, "Disabled": GR_DISABLED}

EFFECT_TIDES = 1 << 0
EFFECT_ROTATIONAL_FLATTENING = 1 << 1
EFFECT_GENERAL_RELATIVITY = 1 << 2
EFFECT_DISK = 1 << 3
EFFECT_WIND = 1 << 4
EFFECT_EVOLUTION = 1 << 5

ROLE_TIDES_CENTRAL = 1 << 0
ROLE_TIDES_ORBITING = 1 << 1
ROLE_FLAT_CENTRAL = 1 << 2
ROLE_FLAT_ORBITING = 1 << 3
ROLE_GR_CENTRAL = 1 << 4
ROLE_GR_ORBITING = 1 << 5
ROLE_WIND = 1 << 6
ROLE_DISK_CENTRAL = 1 << 7
ROLE_DISK_ORBITING = 1 << 8

EVO_NON_EVOLVING = 0
EVO_LECONTE2011 = 1
EVO_BARAFFE1998 = 2
EVO_BARAFFE2015 = 3
EVO_LECONTE_CHABRIER2013 = 4
EVO_LECONTE_CHABRIER2013_DYN = 5
EVO_BOLMONT_MATHIS2016 = 6
EVO_GALLET_BOLMONT2017 = 7

STATUS_COMPLETED = 1 << 0
STATUS_ROCHE = 1 << 1
STATUS_COLLISION = 1 << 2
STATUS_EJECTION = 1 << 3
STATUS_ZERO_INERTIA = 1 << 4
STATUS_MIDPOINT_NOT_CONVERGED = 1 << 5
STATUS_TIMESTEP_WARNING = 1 << 6
STATUS_KEPLER_BISECTION = 1 << 7
STATUS_ITERATION_CAP = 1 << 8
STATUS_HISTORY_OVERFLOW = 1 << 9
STATUS_FATAL = STATUS_ROCHE | STATUS_COLLISION | STATUS_EJECTION | STATUS_ZERO_INERTIA | STATUS_ITERATION_CAP

NOALLOC_MAX_BODIES = 524288

MAPPING_AUTO = 0
MAPPING_PLAIN = 1
MAPPING_PACKED = 2

_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)
_pu32 = C.POINTER(C.c_uint32)
_pi64 = C.POINTER(C.c_int64)


class EnsembleDesc(C.Structure):
    _fields_ = [
        ("n_systems", C.c_int32),
        ("body_offset", _pi32),
        ("coordinates", _pi32),
        ("effects", _pu32),
        ("gr_implementation", _pi32),
        ("time_step", _pd),
        ("time_limit", _pd),
        ("current_time", _pd),
        ("historic_snapshot_period", _pd),
        ("last_historic_snapshot_time", _pd),
        ("current_iteration", _pi64),
        ("mass", _pd),
        ("radius", _pd),
        ("radius_of_gyration_2", _pd),
        ("moment_of_inertia", _pd),
        ("position", _pd),
        ("velocity", _pd),
        ("spin", _pd),
        ("angular_momentum", _pd),
        ("heliocentric_velocity", _pd),
        ("roles", _pu32),
        ("tides_dissipation_factor", _pd),
        ("tides_dissipation_factor_scale", _pd),
        ("tides_love_number", _pd),
        ("flattening_love_number", _pd),
        ("wind_k_factor", _pd),
        ("wind_rotation_saturation", _pd),
        ("velocity_error", _pd),
        ("angular_momentum_error", _pd),
        ("evolution_type", _pi32),
        ("evolver_offset", _pi32),
        ("evolver_time", _pd),
        ("evolver_radius", _pd),
        ("evolver_radius_of_gyration_2", _pd),
        ("evolver_love_number", _pd),
        ("evolver_left_index", _pi32),
        ("evolver_inverse_tidal_q_factor", _pd),
        ("disk_properties", _pd),
        ("pair_dissipation_factor", _pd),
        ("roche_radius", _pd),
    ]


class State(C.Structure):
    _fields_ = [
        ("position", _pd),
        ("velocity", _pd),
        ("acceleration", _pd),
        ("angular_momentum", _pd),
        ("spin", _pd),
        ("velocity_error", _pd),
        ("angular_momentum_error", _pd),
        ("radius", _pd),
        ("radius_of_gyration_2", _pd),
        ("moment_of_inertia", _pd),
        ("current_time", _pd),
        ("last_historic_snapshot_time", _pd),
        ("current_iteration", _pi64),
        ("status", _pu32),
        ("evolver_left_index", _pi32),
        ("timestep_warning", _pi32),
        ("pair_dissipation_factor", _pd),
        ("roche_radius", _pd),
    ]


class Options(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("mapping", C.c_int32),
        ("group_size", C.c_int32),
        ("steps_per_launch", C.c_int32),
        ("use_graph", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


class LaunchStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("kernel_launches", C.c_int64), ("graph_replays", C.c_int64),
                ("launches_per_replay", C.c_int64), ("last_operator_kernel_ms", C.c_double)]


class HistoryRecord(C.Structure):
    _fields_ = [("system", C.c_int32), ("body", C.c_int32), ("time", C.c_double), ("time_step", C.c_double),
                ("position", C.c_double * 3), ("spin", C.c_double * 3), ("velocity", C.c_double * 3),
                ("mass", C.c_double), ("radius", C.c_double), ("radius_of_gyration_2", C.c_double),
                ("love_number", C.c_double), ("scaled_dissipation_factor", C.c_double), ("lag_angle", C.c_double),
                ("denergy_dt", C.c_double), ("migration_timescale", C.c_double)]


class AdjointSeed(C.Structure):
    _fields_ = [("position", _pd), ("velocity", _pd), ("angular_momentum", _pd)]


class AdjointGrad(C.Structure):
    _fields_ = [("position", _pd), ("velocity", _pd), ("angular_momentum", _pd),
                ("dissipation_factor", _pd), ("love_number", _pd), ("fluid_love_number", _pd)]


def ptr(a, ctype):
    """Pointer to the data of a contiguous numpy array, or NULL for None."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "arrays handed to the C ABI must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


STATE_FIELDS = {
    # name: (per-body width, dtype) ; width 0 = per system
    "position": (3, np.float64),
    "velocity": (3, np.float64),
    "acceleration": (3, np.float64),
    "angular_momentum": (3, np.float64),
    "spin": (3, np.float64),
    "velocity_error": (3, np.float64),
    "angular_momentum_error": (3, np.float64),
    "radius": (1, np.float64),
    "radius_of_gyration_2": (1, np.float64),
    "moment_of_inertia": (1, np.float64),
    "current_time": (0, np.float64),
    "last_historic_snapshot_time": (0, np.float64),
    "current_iteration": (0, np.int64),
    "status": (0, np.uint32),
    "evolver_left_index": (1, np.int32),
    "timestep_warning": (0, np.int32),
    "pair_dissipation_factor": (1, np.float64),
    "roche_radius": (MAX_BODIES, np.float64),
}

_CTYPE = {np.float64: C.c_double, np.int64: C.c_int64, np.uint32: C.c_uint32, np.int32: C.c_int32}


def alloc_state(n_systems: int, n_bodies: int, fields=None):
    """numpy buffers for a pos_state plus the filled ctypes struct."""
    arrays = {}
    st = State()
    for name, (w, dt) in STATE_FIELDS.items():
        if fields is not None and name not in fields:
            continue
        if w == 0:
            a = np.zeros(n_systems, dtype=dt)
        elif w == 1:
            a = np.zeros(n_bodies, dtype=dt)
        else:
            a = np.zeros((n_bodies, w), dtype=dt)
        arrays[name] = a
        setattr(st, name, ptr(a, _CTYPE[dt]))
    return st, arrays


def state_struct(arrays: dict):
    """ctypes pos_state pointing at existing numpy arrays (missing -> NULL)."""
    st = State()
    for name, (w, dt) in STATE_FIELDS.items():
        a = arrays.get(name)
        if a is not None:
            a = np.ascontiguousarray(a, dtype=dt)
            arrays[name] = a
        setattr(st, name, ptr(a, _CTYPE[dt]))
    return st


# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libposidonius_hip.so"
LIB_PATH = os.environ.get("POSIDONIUS_HIP_LIBRARY") or os.path.join(_HERE, LIB_NAME)
_lib = None


class PosidoniusError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"posidonius_hip error {code}: {message}")
        self.code = code


def _declare(lib):
    pe = C.POINTER(C.c_void_p)
    lib.pos_abi_version.restype = C.c_int32
    lib.pos_last_error.restype = C.c_char_p
    lib.pos_create.argtypes = [C.POINTER(EnsembleDesc), C.POINTER(Options), pe]
    lib.pos_create.restype = C.c_int32
    lib.pos_step.argtypes = [C.c_void_p, C.c_int64]
    lib.pos_step.restype = C.c_int32
    lib.pos_synchronize.argtypes = [C.c_void_p]
    lib.pos_synchronize.restype = C.c_int32
    lib.pos_get_state.argtypes = [C.c_void_p, C.POINTER(State)]
    lib.pos_get_state.restype = C.c_int32
    lib.pos_set_state.argtypes = [C.c_void_p, C.POINTER(State)]
    lib.pos_set_state.restype = C.c_int32
    lib.pos_set_time_limit.argtypes = [C.c_void_p, C.c_double]
    lib.pos_set_time_limit.restype = C.c_int32
    lib.pos_set_historic_snapshot_period.argtypes = [C.c_void_p, C.c_double]
    lib.pos_set_historic_snapshot_period.restype = C.c_int32
    lib.pos_get_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.pos_get_counts.restype = C.c_int32
    lib.pos_get_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    lib.pos_get_stream.restype = C.c_int32
    lib.pos_get_options.argtypes = [C.c_void_p, C.POINTER(Options)]
    lib.pos_get_options.restype = C.c_int32
    lib.pos_get_launch_stats.argtypes = [C.c_void_p, C.POINTER(LaunchStats)]
    lib.pos_get_launch_stats.restype = C.c_int32
    lib.pos_tune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Options)]
    lib.pos_tune.restype = C.c_int32
    lib.pos_gravity_acceleration.argtypes = [C.c_void_p, _pd, _pu32]
    lib.pos_gravity_acceleration.restype = C.c_int32
    lib.pos_additional_effects.argtypes = [C.c_void_p, _pd, _pd]
    lib.pos_additional_effects.restype = C.c_int32
    lib.pos_additional_effects_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.pos_additional_effects_noalloc.restype = C.c_int32
    lib.pos_history_enable.argtypes = [C.c_void_p, C.c_int32]
    lib.pos_history_enable.restype = C.c_int32
    lib.pos_history_drain.argtypes = [C.c_void_p, C.POINTER(HistoryRecord), C.c_int64, C.POINTER(C.c_int64)]
    lib.pos_history_drain.restype = C.c_int32
    lib.pos_adjoint_record.argtypes = [C.c_void_p, C.c_int64, C.c_int32]
    lib.pos_adjoint_record.restype = C.c_int32
    lib.pos_adjoint.argtypes = [C.c_void_p, C.POINTER(AdjointSeed), C.POINTER(AdjointGrad)]
    lib.pos_adjoint.restype = C.c_int32
    lib.pos_destroy.argtypes = [C.c_void_p]
    lib.pos_destroy.restype = None


EXPORTED_SYMBOLS = (
    "pos_abi_version", "pos_last_error", "pos_create", "pos_step", "pos_synchronize",
    "pos_get_state", "pos_set_state", "pos_set_time_limit", "pos_set_historic_snapshot_period",
    "pos_get_counts", "pos_get_stream", "pos_get_options", "pos_get_launch_stats", "pos_tune",
    "pos_gravity_acceleration",
    "pos_additional_effects", "pos_additional_effects_noalloc", "pos_history_enable", "pos_history_drain", "pos_adjoint_record", "pos_adjoint",
    "pos_destroy",
)


def library():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the integration step has no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        _declare(lib)
        v = lib.pos_abi_version()
        if v != ABI_VERSION:
            raise ImportError(f"{LIB_PATH} has ABI version {v}, expected {ABI_VERSION}")
        _lib = lib
    return _lib


def check(code: int):
    if code != OK:
        msg = library().pos_last_error()
        raise PosidoniusError(code, msg.decode() if msg else "")
