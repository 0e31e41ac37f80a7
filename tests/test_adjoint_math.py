"""Derivative checks of the per-lane adjoint math (csrc/adjoint_math.hpp).

The header is compiled for the host (tests/cpu/adjoint_math_capi.cpp) and
every vector-Jacobian product is compared with central finite differences of
the forward function it belongs to, one input block and one output block at a
time, so that a wrong term in any single block fails on its own.  The pair
forward function is also checked against the oracle's additional-effects pass
(the reference's equations), so the adjoint linearises the right function.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from posidonius_amd import abi, generators as gen

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "adjoint_math_capi.cpp")
HDR = os.path.join(os.path.dirname(HERE), "posidonius_amd", "csrc", "adjoint_math.hpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    # POSIDONIUS_ADJCHECK_LIBRARY: a prebuilt library of the same source (the
    # sanitizer build of tests/cpu/Makefile)
    out = os.environ.get("POSIDONIUS_ADJCHECK_LIBRARY")
    if not out:
        out = str(tmp_path_factory.mktemp("adj") / "libadjcheck.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", SRC, "-o", out], check=True)
    lib = C.CDLL(out)
    d = C.POINTER(C.c_double)
    lib.adj_pair_forward.argtypes = [d, d, d, C.c_int, d]
    lib.adj_pair_vjp.argtypes = [d, d, d, C.c_int, d, d, d]
    lib.adj_stiefel.argtypes = [C.c_double, C.c_double, d]
    lib.adj_kepler_residual.argtypes = [C.c_double, C.c_double, C.c_double, d, d]
    lib.adj_kepler_residual.restype = C.c_double
    lib.adj_kepler_forward.argtypes = [C.c_double, C.c_double, C.c_double, d, d, d, d]
    lib.adj_kepler_vjp.argtypes = [C.c_double, C.c_double, C.c_double, d, d, d, d, d, d]
    lib.adj_gravity_pair_vjp.argtypes = [C.c_double, d, d, d]
    return lib


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def pair_case(spec, s=0, planet=1):
    b0 = spec.body_offset[s]
    h, p = b0, b0 + planet
    B = spec.body
    x, v, W = B["position"], B["velocity"], B["spin"]
    inp = np.concatenate([x[p] - x[h], v[p] - v[h], W[h], W[p]])
    I = B["mass"] * B["radius_of_gyration_2"] * B["radius"] ** 2
    cst = np.array([B["mass"][h], B["mass"][p], B["radius"][h], B["radius"][p], I[h], I[p]])
    sig = B["tides_dissipation_factor"] * B["tides_dissipation_factor_scale"]
    par = np.array([sig[h], sig[p], B["tides_love_number"][h], B["tides_love_number"][p],
                    B["flattening_love_number"][h], B["flattening_love_number"][p]])
    return inp, cst, par


def fwd(lib, inp, cst, par, flags):
    out = np.zeros(12)
    lib.adj_pair_forward(P(inp), P(cst), P(par), flags, P(out))
    return out


def close(a, b, rtol, floor=0.0):
    """|a - b| within rtol of the larger, plus an absolute floor (the
    rounding noise of the finite difference)."""
    return abs(a - b) <= rtol * max(abs(a), abs(b)) + floor + 1e-300


def fd_noise(ob, out, h):
    """Rounding noise of a central difference of ob . out with step h."""
    return 20 * np.finfo(float).eps * np.abs(ob) @ np.abs(out) / h


@pytest.mark.parametrize("flags", [1, 2, 4, 7])
def test_pair_vjp_matches_finite_differences(lib, flags):
    spec = gen.config4(4)
    rng = np.random.default_rng(11 + flags)
    inp, cst, par = pair_case(spec)
    base = fwd(lib, inp, cst, par, flags)
    for ob_block in range(4):
        ob = np.zeros(12)
        ob[3 * ob_block:3 * ob_block + 3] = rng.standard_normal(3)
        if not np.any(base[3 * ob_block:3 * ob_block + 3]):
            continue          # this effect does not produce that output
        ib = np.zeros(12)
        pb = np.zeros(6)
        lib.adj_pair_vjp(P(inp), P(cst), P(par), flags, P(ob), P(ib), P(pb))
        for in_block in range(4):
            seg = slice(3 * in_block, 3 * in_block + 3)
            delta = np.zeros(12)
            delta[seg] = rng.standard_normal(3) * np.linalg.norm(inp[seg])
            h = 1e-6
            fd = (ob @ fwd(lib, inp + h * delta, cst, par, flags) - ob @ fwd(lib, inp - h * delta, cst, par, flags)) / (2 * h)
            adj = ib @ delta
            assert close(fd, adj, 2e-6, fd_noise(ob, base, h)), (flags, ob_block, in_block, fd, adj)
        for k in range(6):
            if par[k] == 0:
                continue
            dp = np.zeros(6)
            dp[k] = par[k]
            h = 1e-6
            fd = (ob @ fwd(lib, inp, cst, par + h * dp, flags) - ob @ fwd(lib, inp, cst, par - h * dp, flags)) / (2 * h)
            assert close(fd, pb[k] * par[k], 2e-6, fd_noise(ob, base, h)), (flags, ob_block, "param", k, fd, pb[k] * par[k])


def test_pair_forward_matches_oracle_effects(lib, oracle):
    """Sum of pair_forward over the planets equals the oracle's additional
    accelerations and dL/dt (to rounding) for a star + 3 planet system with
    tides, flattening and Kidder GR."""
    spec = gen.config4(2)
    acc, dl = oracle.additional_effects(spec)
    for s in range(spec.n_systems):
        b0, b1 = spec.body_offset[s], spec.body_offset[s + 1]
        m = spec.body["mass"]
        a_h = np.zeros(3)
        dl_h = np.zeros(3)
        for p in range(1, b1 - b0):
            inp, cst, par = pair_case(spec, s, p)
            o = fwd(lib, inp, cst, par, 7)
            F, g, Np, Nh = o[:3], o[3:6], o[6:9], o[9:]
            ap = F / m[b0 + p] + g
            np.testing.assert_allclose(ap, acc[b0 + p], rtol=1e-11, atol=1e-13 * np.abs(acc[b0 + p]).max())
            np.testing.assert_allclose(Np, dl[b0 + p], rtol=1e-11, atol=1e-13 * np.abs(dl[b0 + p]).max())
            a_h += -F / m[b0] - (m[b0 + p] / m[b0]) * g
            dl_h += Nh
        np.testing.assert_allclose(a_h, acc[b0], rtol=1e-10, atol=1e-12 * np.abs(acc[b0]).max())
        np.testing.assert_allclose(dl_h, dl[b0], rtol=1e-10, atol=1e-12 * np.abs(dl[b0]).max())


def test_stiefel_functions_closed_forms_and_beta_derivative(lib):
    g = np.zeros(6)
    for beta, X in [(3.0, 0.2), (0.5, 2.0), (-0.7, 0.9), (2.0, 1.3), (1e-3, 0.1)]:
        lib.adj_stiefel(beta, X, P(g))
        s = np.sqrt(abs(beta))
        if beta > 0:
            c0, c1 = np.cos(s * X), np.sin(s * X) / s
        else:
            c0, c1 = np.cosh(s * X), np.sinh(s * X) / s
        np.testing.assert_allclose(g[0], c0, rtol=1e-13)
        np.testing.assert_allclose(g[1], c1, rtol=1e-13)
        np.testing.assert_allclose(g[2], (1 - c0) / beta, rtol=1e-9)
        np.testing.assert_allclose(g[3], (X - c1) / beta, rtol=1e-9)
        # dG_k/dbeta = (k G_{k+2} - X G_{k+1}) / 2 and dG_k/dX = G_{k-1}
        gp, gm = np.zeros(6), np.zeros(6)
        h = 1e-4 * abs(beta)
        lib.adj_stiefel(beta + h, X, P(gp))
        lib.adj_stiefel(beta - h, X, P(gm))
        for k in (1, 2, 3):
            np.testing.assert_allclose((gp[k] - gm[k]) / (2 * h), (k * g[k + 2] - X * g[k + 1]) / 2, rtol=1e-5)
        hx = 1e-6 * X
        lib.adj_stiefel(beta, X + hx, P(gp))
        lib.adj_stiefel(beta, X - hx, P(gm))
        for k in (1, 2, 3, 4, 5):
            np.testing.assert_allclose((gp[k] - gm[k]) / (2 * hx), g[k - 1], rtol=1e-6)


def solve_x(lib, mu, dt, x0, v0):
    r0 = np.linalg.norm(x0)
    X = dt / r0
    for _ in range(60):
        f = lib.adj_kepler_residual(mu, dt, X, P(x0), P(v0))
        g = np.zeros(6)
        beta = 2 * mu / r0 - v0 @ v0
        lib.adj_stiefel(beta, X, P(g))
        fp = r0 + (x0 @ v0) * g[1] + (mu - beta * r0) * g[2]
        X -= f / fp
    return X


def kep(lib, mu, dt, x0, v0):
    X = solve_x(lib, mu, dt, x0, v0)
    x1, v1 = np.zeros(3), np.zeros(3)
    lib.adj_kepler_forward(mu, dt, X, P(x0), P(v0), P(x1), P(v1))
    return x1, v1, X


def test_kepler_drift_matches_two_body_closed_form(lib):
    """Circular orbit: the drift is a rotation by n dt."""
    mu = 0.09 * 2.959122082855911e-4
    a = 0.012
    n = np.sqrt(mu / a ** 3)
    x0, v0 = np.array([a, 0, 0.]), np.array([0, n * a, 0.])
    dt = 0.3
    x1, v1, _ = kep(lib, mu, dt, x0, v0)
    np.testing.assert_allclose(x1, [a * np.cos(n * dt), a * np.sin(n * dt), 0], atol=1e-15)
    np.testing.assert_allclose(v1, [-n * a * np.sin(n * dt), n * a * np.cos(n * dt), 0], atol=1e-15 * n * a / a)


def test_kepler_vjp_matches_finite_differences(lib):
    mu = 0.09 * 2.959122082855911e-4
    rng = np.random.default_rng(5)
    for trial in range(6):
        a = 0.01 + 0.01 * rng.random()
        e = 0.3 * rng.random()
        x0 = np.array([a * (1 - e), 0.001 * rng.standard_normal(), 0.0005 * rng.standard_normal()])
        v0 = np.array([0.002 * rng.standard_normal(), np.sqrt(mu * (1 + e) / (a * (1 - e))), 0.001 * rng.standard_normal()])
        dt = 0.05 + 0.5 * rng.random()
        x1, v1, X = kep(lib, mu, dt, x0, v0)
        xb, vb = rng.standard_normal(3), rng.standard_normal(3) / np.linalg.norm(v0) * np.linalg.norm(x0)
        gx, gv = np.zeros(3), np.zeros(3)
        lib.adj_kepler_vjp(mu, dt, X, P(x0), P(v0), P(xb), P(vb), P(gx), P(gv))
        for block in range(2):
            d = rng.standard_normal(3) * (np.linalg.norm(x0) if block == 0 else np.linalg.norm(v0))
            h = 1e-6
            if block == 0:
                p1, q1, _ = kep(lib, mu, dt, x0 + h * d, v0)
                p2, q2, _ = kep(lib, mu, dt, x0 - h * d, v0)
                adj = gx @ d
            else:
                p1, q1, _ = kep(lib, mu, dt, x0, v0 + h * d)
                p2, q2, _ = kep(lib, mu, dt, x0, v0 - h * d)
                adj = gv @ d
            fd = (xb @ (p1 - p2) + vb @ (q1 - q2)) / (2 * h)
            assert close(fd, adj, 1e-6), (trial, block, fd, adj)


def test_gravity_pair_vjp_matches_finite_differences(lib):
    rng = np.random.default_rng(3)
    AU, DAY = 1.49597870700e11, 86400.
    G = 6.67428e-11 / AU ** 3 * 1.9818e30 * DAY ** 2     # the reference's K2 (src/constants.rs)
    for _ in range(5):
        d = rng.standard_normal(3) * 0.01
        ab = rng.standard_normal(3)
        mj = 3e-6
        out = np.zeros(3)
        lib.adj_gravity_pair_vjp(mj, P(d), P(ab), P(out))
        acc = lambda dd: -G * mj * dd / np.linalg.norm(dd) ** 3
        delta = rng.standard_normal(3) * 0.01
        h = 1e-6
        fd = ab @ (acc(d + h * delta) - acc(d - h * delta)) / (2 * h)
        assert close(fd, out @ delta, 1e-7)
