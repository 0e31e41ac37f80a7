"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is built from oracle/ by
__graft_entry__.build() (or `make -C oracle`) and loaded only here, in
smoke() and in bench.py's cpu_baseline leg.
"""
import ctypes as C
import glob
import json
import os
import subprocess

import numpy as np

from posidonius_amd import abi, case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# POSIDONIUS_ORACLE_LIBRARY selects another build of the same sources (the
# sanitizer build of tests/cpu/Makefile)
LIB = os.environ.get("POSIDONIUS_ORACLE_LIBRARY") or os.path.join(ORACLE_DIR, "liboracle.so")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class Oracle:
    def __init__(self):
        if not os.path.exists(LIB):
            build_oracle()
        self.lib = C.CDLL(LIB)
        self.lib.oracle_last_error.restype = C.c_char_p
        self.lib.oracle_run.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        self.lib.oracle_gravity.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.lib.oracle_additional_effects.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.lib.oracle_flops.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        self.lib.oracle_tangent.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32] + [C.c_void_p] * 10

    def _err(self):
        return self.lib.oracle_last_error().decode()

    def run(self, spec, n_steps, threads=1, fields=None):
        h = spec.desc()
        st, arr = abi.alloc_state(spec.n_systems, spec.n_bodies, fields)
        steps = np.zeros(spec.n_systems, dtype=np.int64)
        mid = np.zeros(spec.n_systems, dtype=np.int64)
        rc = self.lib.oracle_run(C.addressof(h.desc), int(n_steps), int(threads), C.addressof(st),
                                 steps.ctypes.data, mid.ctypes.data)
        if rc != 0:
            raise RuntimeError(self._err())
        arr["steps_done"] = steps
        arr["midpoint_evaluations"] = mid
        return arr

    def flops(self, spec, n_steps):
        """FP64 operations (+ - * / sqrt, each counted once) of n_steps
        iterations of every system, counted by the oracle's FlopCount scalar."""
        h = spec.desc()
        out = np.zeros(spec.n_systems)
        steps = np.zeros(spec.n_systems, dtype=np.int64)
        if self.lib.oracle_flops(C.addressof(h.desc), int(n_steps), out.ctypes.data, steps.ctypes.data):
            raise RuntimeError(self._err())
        return out, steps

    def tangent(self, spec, n_steps, dirs, threads=8):
        """Tangent-linear run: `dirs` maps any of x, v, L ([K, nb, 3]) and
        sigma, k2, k2f ([K, nb]) to K seed directions (K <= 8).  Returns the
        final state and the tangents of x, v, L, each [K, nb, 3]."""
        K = next(iter(dirs.values())).shape[0]
        nb = spec.n_bodies
        h = spec.desc()
        st, arr = abi.alloc_state(spec.n_systems, nb)
        seeds = {}
        for k in ("x", "v", "L", "sigma", "k2", "k2f"):
            a = dirs.get(k)
            seeds[k] = None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        outs = {k: np.zeros((K, nb, 3)) for k in ("x", "v", "L")}
        ptr = lambda a: None if a is None else a.ctypes.data
        rc = self.lib.oracle_tangent(C.addressof(h.desc), int(n_steps), int(threads), int(K),
                                     ptr(seeds["x"]), ptr(seeds["v"]), ptr(seeds["L"]), ptr(seeds["sigma"]),
                                     ptr(seeds["k2"]), ptr(seeds["k2f"]), C.addressof(st),
                                     outs["x"].ctypes.data, outs["v"].ctypes.data, outs["L"].ctypes.data)
        if rc != 0:
            raise RuntimeError(self._err())
        return arr, outs

    def gravity(self, spec):
        h = spec.desc()
        acc = np.zeros((spec.n_bodies, 3))
        status = np.zeros(spec.n_systems, dtype=np.uint32)
        if self.lib.oracle_gravity(C.addressof(h.desc), acc.ctypes.data, status.ctypes.data):
            raise RuntimeError(self._err())
        return acc, status

    def additional_effects(self, spec):
        h = spec.desc()
        acc = np.zeros((spec.n_bodies, 3))
        dl = np.zeros((spec.n_bodies, 3))
        if self.lib.oracle_additional_effects(C.addressof(h.desc), acc.ctypes.data, dl.ctypes.data):
            raise RuntimeError(self._err())
        return acc, dl


def golden_names():
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN_DIR, "test_*.json")))


def load_golden(name):
    with open(os.path.join(GOLDEN_DIR, name + ".json")) as f:
        g = json.load(f)
    spec = case.read_case(g["case"])
    expected = {
        "position": np.array([e["inertial_position"] for e in g["expected"]]),
        "velocity": np.array([e["inertial_velocity"] for e in g["expected"]]),
        "acceleration": np.array([e["inertial_acceleration"] for e in g["expected"]]),
    }
    return spec, expected
