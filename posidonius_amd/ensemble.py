"""Python handle over the C ABI: an ensemble of WHFast systems on one GPU.

The object mirrors the reference's ``Integrator`` trait
(src/integrator/mod.rs:16-26) for a batch of universes: ``step`` is
``iterate`` applied ``n`` times to every system, ``set_time_limit`` and
``set_historic_snapshot_period`` are the trait's setters, and the state
accessors expose what ``Universe.particles`` holds.  All work happens in
``libposidonius_hip.so``; if the library is absent, constructing an
``Ensemble`` raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import abi
from .case import EnsembleSpec

MAPPINGS = {"auto": abi.MAPPING_AUTO, "plain": abi.MAPPING_PLAIN, "packed": abi.MAPPING_PACKED}


class Ensemble:
    def __init__(self, spec: EnsembleSpec, mapping: str = "packed", group_size: int = 0,
                 steps_per_launch: int = 0, use_graph: bool = False, device: int = 0,
                 persistent_grid: int = 0):
        self._lib = abi.library()
        self.spec = spec
        self._holder = spec.desc()
        opts = abi.Options()
        opts.device = device
        opts.mapping = MAPPINGS[mapping] if isinstance(mapping, str) else int(mapping)
        opts.group_size = group_size
        opts.steps_per_launch = steps_per_launch
        opts.use_graph = 1 if use_graph else 0
        opts.reserved[0] = persistent_grid
        h = C.c_void_p()
        abi.check(self._lib.pos_create(self._holder.ref, C.byref(opts), C.byref(h)))
        self._h = h
        self.n_systems = spec.n_systems
        self.n_bodies = spec.n_bodies

    # -- lifecycle ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.pos_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping -----------------------------------------------------------
    def step(self, n_steps: int = 1):
        abi.check(self._lib.pos_step(self._h, C.c_int64(int(n_steps))))

    def synchronize(self):
        abi.check(self._lib.pos_synchronize(self._h))

    def tune(self, probe_steps: int = 32) -> abi.Options:
        o = abi.Options()
        abi.check(self._lib.pos_tune(self._h, int(probe_steps), C.byref(o)))
        return o

    def stream(self) -> int:
        """hipStream_t of the handle (as an integer address)."""
        st = C.c_void_p()
        abi.check(self._lib.pos_get_stream(self._h, C.byref(st)))
        return st.value or 0

    def options(self) -> abi.Options:
        o = abi.Options()
        abi.check(self._lib.pos_get_options(self._h, C.byref(o)))
        return o

    def launch_stats(self) -> Dict[str, float]:
        """What the handle has executed so far (pos_get_launch_stats)."""
        st = abi.LaunchStats()
        abi.check(self._lib.pos_get_launch_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in abi.LaunchStats._fields_}

    def set_time_limit(self, time_limit: float):
        abi.check(self._lib.pos_set_time_limit(self._h, float(time_limit)))

    def set_historic_snapshot_period(self, period: float):
        abi.check(self._lib.pos_set_historic_snapshot_period(self._h, float(period)))

    # -- state ----------------------------------------------------------------
    def get_state(self, fields=None) -> Dict[str, np.ndarray]:
        st, arrays = abi.alloc_state(self.n_systems, self.n_bodies, fields)
        abi.check(self._lib.pos_get_state(self._h, C.byref(st)))
        return arrays

    def set_state(self, arrays: Dict[str, np.ndarray]):
        arrays = dict(arrays)
        st = abi.state_struct(arrays)
        abi.check(self._lib.pos_set_state(self._h, C.byref(st)))

    # -- operator-level evaluations ------------------------------------------
    def gravity_acceleration(self):
        acc = np.zeros((self.n_bodies, 3))
        status = np.zeros(self.n_systems, dtype=np.uint32)
        abi.check(self._lib.pos_gravity_acceleration(self._h, abi.ptr(acc, C.c_double), abi.ptr(status, C.c_uint32)))
        return acc, status

    def additional_effects(self):
        acc = np.zeros((self.n_bodies, 3))
        dl = np.zeros((self.n_bodies, 3))
        abi.check(self._lib.pos_additional_effects(self._h, abi.ptr(acc, C.c_double), abi.ptr(dl, C.c_double)))
        return acc, dl

    # -- historic snapshots ---------------------------------------------------
    def enable_history(self, capacity: int):
        """Keep up to `capacity` historic-snapshot records per system on the
        device between drains."""
        abi.check(self._lib.pos_history_enable(self._h, int(capacity)))
        self._hist_cap = int(capacity)

    def drain_history(self) -> np.ndarray:
        """Pending historic-snapshot records (structured array, input order:
        system, record, body); clears the device buffer."""
        cap = getattr(self, "_hist_cap", 0)
        n_max = cap * self.n_bodies
        buf = (abi.HistoryRecord * max(n_max, 1))()
        n = C.c_int64()
        abi.check(self._lib.pos_history_drain(self._h, buf, n_max, C.byref(n)))
        dt = np.dtype([("system", np.int32), ("body", np.int32), ("time", np.float64), ("time_step", np.float64),
                       ("position", np.float64, 3), ("spin", np.float64, 3), ("velocity", np.float64, 3),
                       ("mass", np.float64), ("radius", np.float64), ("radius_of_gyration_2", np.float64),
                       ("love_number", np.float64), ("scaled_dissipation_factor", np.float64),
                       ("lag_angle", np.float64), ("denergy_dt", np.float64), ("migration_timescale", np.float64)])
        assert dt.itemsize == C.sizeof(abi.HistoryRecord)
        return np.frombuffer(bytes(buf)[:n.value * dt.itemsize], dtype=dt).copy()

    # -- discrete adjoint -----------------------------------------------------
    def adjoint_record(self, n_steps: int, segment_steps: int = 0):
        """Advance n_steps like ``step`` and keep a checkpoint every
        ``segment_steps`` steps (0 = 1000) for ``adjoint``."""
        abi.check(self._lib.pos_adjoint_record(self._h, C.c_int64(int(n_steps)), int(segment_steps)))

    def adjoint(self, position=None, velocity=None, angular_momentum=None) -> Dict[str, np.ndarray]:
        """Vector-Jacobian product of the recorded run.  The arguments are the
        cotangents of the final position, velocity and spin angular momentum
        ([n_bodies, 3], None = 0).  Returns the cotangents of the initial
        state and the gradients with respect to every body's dissipation
        factor sigma (scale * factor), tidal Love number and fluid Love number."""
        nb = self.n_bodies
        keep = []

        def src(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_double))
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(nb, 3)
            keep.append(a)
            return abi.ptr(a, C.c_double)

        seed = abi.AdjointSeed(src(position), src(velocity), src(angular_momentum))
        out = {"position": np.zeros((nb, 3)), "velocity": np.zeros((nb, 3)), "angular_momentum": np.zeros((nb, 3)),
               "dissipation_factor": np.zeros(nb), "love_number": np.zeros(nb), "fluid_love_number": np.zeros(nb)}
        grad = abi.AdjointGrad(*(abi.ptr(out[k], C.c_double) for k in
                                 ("position", "velocity", "angular_momentum", "dissipation_factor", "love_number",
                                  "fluid_love_number")))
        abi.check(self._lib.pos_adjoint(self._h, C.byref(seed), C.byref(grad)))
        return out
