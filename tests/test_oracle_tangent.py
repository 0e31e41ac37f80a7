"""The oracle's tangent-linear mode (the checker of the GPU adjoint) against
central finite differences of the oracle itself.

One step of star + test-particle systems whose Kepler drift falls back to
bisection (whfast.rs:776-808): the step's map follows the root of the
universal Kepler equation, whose derivative is the implicit-function one.
Differentiating the bisection iterates literally would give the derivatives
of the bracket's end points instead, off by orders of magnitude; the
tangent-linear mode takes the root's derivative, which the finite
differences confirm.  Newton-path systems are checked alongside.
"""
import numpy as np

from posidonius_amd import abi
from tests.test_gpu_kepler_paths import GM, two_body_ensemble
from tests.test_oracle_kepler import bracket_overflows


def test_tangent_of_the_bisection_root_matches_finite_differences(oracle):
    spec = two_body_ensemble(60, 2026)
    keep = []
    for s in range(spec.n_systems):
        b = spec.body_offset[s] + 1
        x, v = spec.body["position"][b] - spec.body["position"][b - 1], spec.body["velocity"][b] - spec.body["velocity"][b - 1]
        if s % 3 == 2:   # elliptic steps of many orbits: Laguerre-Conway; not needed here
            continue
        if x @ v < 0 and bracket_overflows(GM, spec.system["time_step"][s], x, v):
            continue     # the reference's NaN-bracket quirk ends away from the root
        keep.append(s)
    spec = spec.subset(np.array(keep))
    nb = spec.n_bodies
    rng = np.random.default_rng(4)
    d = {"x": np.zeros((2, nb, 3)), "v": np.zeros((2, nb, 3))}
    scale_x = np.abs(spec.body["position"]).max(axis=1, keepdims=True)
    scale_v = np.abs(spec.body["velocity"]).max(axis=1, keepdims=True)
    d["x"][0] = rng.standard_normal((nb, 3)) * scale_x
    d["v"][1] = rng.standard_normal((nb, 3)) * scale_v
    state, tan = oracle.tangent(spec, 1, d, threads=4)
    assert np.count_nonzero(state["status"] & abi.STATUS_KEPLER_BISECTION) >= 10
    h = 1e-7
    for k, key in ((0, "position"), (1, "velocity")):
        plus, minus = spec.copy(), spec.copy()
        plus.body[key] = spec.body[key] + h * d["x" if k == 0 else "v"][k]
        minus.body[key] = spec.body[key] - h * d["x" if k == 0 else "v"][k]
        rp, rm = oracle.run(plus, 1), oracle.run(minus, 1)
        for out in ("position", "velocity"):
            fd = (rp[out] - rm[out]) / (2 * h)
            td = tan["x" if out == "position" else "v"][k]
            planet = np.arange(nb) % 2 == 1
            err = np.abs(td[planet] - fd[planet]).max(axis=1) / np.abs(fd[planet]).max(axis=1)
            assert err.max() < 1e-4, (key, out, err.max())
