"""Long-horizon behaviour: 10^6 steps of the M-dwarf + two Earths system.

The GPU run must equal the oracle's bit for bit after 10^6 steps, so its
relative energy and angular-momentum drift is the reference algorithm's,
not just of the same magnitude.  The drift itself is reported and bounded:
point-mass energy conserved by the symplectic map to the level of the
dissipative perturbations, total (orbital + spin) angular momentum to
round-off plus the torques' discretisation.
"""
import numpy as np
import pytest

from posidonius_amd import generators as gen

pytestmark = pytest.mark.gpu

G = 6.67428e-11 / 1.49597870700e11 ** 3 * 1.9818e30 * 86400.0 ** 2


def energy_and_momentum(spec, st):
    m = spec.body["mass"]
    x, v, L = st["position"], st["velocity"], st["angular_momentum"]
    kin = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    pot = 0.0
    for i in range(len(m)):
        for j in range(i + 1, len(m)):
            pot -= G * m[i] * m[j] / np.linalg.norm(x[i] - x[j])
    lorb = np.sum(m[:, None] * np.cross(x, v), axis=0)
    return kin + pot, lorb + L.sum(axis=0)


def test_million_steps_equal_oracle_and_drift_is_bounded(oracle):
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config0()
    spec.system["time_limit"][:] = 1e30
    n = 1_000_000
    with Ensemble(spec, mapping="packed", steps_per_launch=1024, use_graph=True) as ens:
        ens.step(n)
        out = ens.get_state()
    ref = oracle.run(spec, n)
    for k in ("position", "velocity", "angular_momentum", "spin", "velocity_error", "angular_momentum_error"):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    assert out["current_iteration"][0] == n
    e0, l0 = energy_and_momentum(spec, {"position": spec.body["position"], "velocity": spec.body["velocity"],
                                        "angular_momentum": spec.body["angular_momentum"]})
    e1, l1 = energy_and_momentum(spec, out)
    de = abs((e1 - e0) / e0)
    dl = np.linalg.norm(l1 - l0) / np.linalg.norm(l0)
    print(f"10^6 steps: relative energy drift {de:.3e}, relative angular-momentum drift {dl:.3e}", flush=True)
    assert de < 1e-6 and dl < 1e-9
