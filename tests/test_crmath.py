"""The correctly rounded exp, x^(-1/2) and x^(-3/2) of the device
(posidonius_amd/csrc/crmath.hpp, compiled here for the host) against
binary128 evaluations (libquadmath, 113-bit significand) rounded to double.

These functions replace f64::exp / f64::powf in the disk and evolving-
dissipation terms (disk.rs:172-226, constant_time_lag.rs:57, 121) on both
the device and the oracle.  The platform libm is reported for comparison:
it is within 0.52 ulp, so a small fraction of its results differ from the
nearest double."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "crmath_capi.cpp")


@pytest.fixture(scope="module")
def lib():
    out = os.path.join(tempfile.mkdtemp(), "crmath_capi.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", SRC, "-o", out,
                    "-lquadmath"], check=True)
    L = ctypes.CDLL(out)
    P = ctypes.POINTER(ctypes.c_double)
    for f in ("cr_exp_n", "cr_rsqrt_n", "cr_rsqrt3_n", "q_exp_n", "libm_exp_n"):
        getattr(L, f).argtypes = [P, ctypes.c_long, P]
    for f in ("q_pow_n", "libm_pow_n"):
        getattr(L, f).argtypes = [P, ctypes.c_long, ctypes.c_double, P]
    return L


def call(L, name, x, *extra):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    P = ctypes.POINTER(ctypes.c_double)
    if extra:
        getattr(L, name)(x.ctypes.data_as(P), x.size, *extra, out.ctypes.data_as(P))
    else:
        getattr(L, name)(x.ctypes.data_as(P), x.size, out.ctypes.data_as(P))
    return out


def report(name, got, want, libm):
    bad = np.flatnonzero(got.view(np.int64) != want.view(np.int64))
    libm_bad = np.count_nonzero(libm.view(np.int64) != want.view(np.int64))
    print(f"{name}: {len(bad)} of {got.size} differ from binary128; libm differs on {libm_bad}")
    return bad


def test_exp_is_correctly_rounded(lib):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-2.0, 2.0, 400_000),          # disk factors exp(-d/r_out), exp(-t/tau)
        rng.uniform(-60.0, 0.0, 300_000),
        rng.uniform(-745.0, 709.7, 300_000),      # whole range, subnormal results included
        np.array([0.0, -0.0, 1e-300, -1e-300, 1e-17, -1e-17, 0.5, -0.5, 1.0, -1.0, 709.78, -708.4, -744.4,
                  -745.13, -745.14, 709.7827128933840]),
    ])
    got = call(lib, "cr_exp_n", x)
    want = call(lib, "q_exp_n", x)
    bad = report("exp", got, want, call(lib, "libm_exp_n", x))
    assert len(bad) == 0, f"first mismatch x={x[bad[0]]!r} got={got[bad[0]]!r} want={want[bad[0]]!r}"
    assert call(lib, "cr_exp_n", np.array([710.0, -800.0, np.inf, -np.inf])).tolist() == [np.inf, 0.0, np.inf, 0.0]
    assert np.isnan(call(lib, "cr_exp_n", np.array([np.nan]))[0])


@pytest.mark.parametrize("fn,y", [("cr_rsqrt_n", -0.5), ("cr_rsqrt3_n", -1.5)])
def test_inverse_powers_are_correctly_rounded(lib, fn, y):
    rng = np.random.default_rng(12)
    x = np.concatenate([
        np.exp(rng.uniform(np.log(1e-4), np.log(1e4), 600_000)),   # distances and semi-major axes in AU
        np.exp(rng.uniform(-400.0, 400.0, 200_000)),
        np.array([1.0, 4.0, 0.25, 2.0, 0.5, 1e-100, 1e100, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0)]),
    ])
    got = call(lib, fn, x)
    want = call(lib, "q_pow_n", x, ctypes.c_double(y))
    bad = report(fn, got, want, call(lib, "libm_pow_n", x, ctypes.c_double(y)))
    assert len(bad) == 0, f"first mismatch x={x[bad[0]]!r} got={got[bad[0]]!r} want={want[bad[0]]!r}"
