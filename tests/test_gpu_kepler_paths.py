"""The device's Kepler drift on the paths no reference fixture reaches:
hyperbolic orbits (Newton and the bisection fallback, whfast.rs:776-808) and
steps longer than an orbit (Laguerre-Conway iterations and the time-step
warning).  Star + test-particle systems with every effect off, so the step
is the drift, the jump and the (empty) interaction; bitwise against the
oracle, whose drift is checked against the closed-form two-body solution in
test_oracle_kepler.py."""
import numpy as np
import pytest

from posidonius_amd import abi, generators as gen
from tests.kepler_closed_form import random_orbit
from tests.test_gpu_parity import assert_same_state, log

pytestmark = pytest.mark.gpu

GM = 2.959122082855911e-04


def two_body_ensemble(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 2, 3))
    vel = np.zeros((n, 2, 3))
    dt = np.zeros(n)
    for s in range(n):
        kind = s % 3
        if kind == 0:     # hyperbolic, long steps: bisection
            x, v = random_orbit(rng, GM, 1.5, 1.95)
            f = 10 ** rng.uniform(0.5, 2.0)
        elif kind == 1:   # hyperbolic, short steps: Newton
            x, v = random_orbit(rng, GM, 1.5, 1.95)
            f = 10 ** rng.uniform(-3.0, -1.0)
        else:             # elliptic, steps of 1-30 orbits: Laguerre-Conway, time-step warning
            x, v = random_orbit(rng, GM, 0.6, 1.3)
            f = rng.uniform(1.05, 30.0) * (2.0 - (v @ v) * np.linalg.norm(x) / GM) ** -1.5
        pos[s, 1], vel[s, 1] = x, v
        dt[s] = f * 2 * np.pi * np.sqrt(np.linalg.norm(x) ** 3 / GM)
    mass = np.tile([1.0, 1e-12], (n, 1))
    radius = np.tile([gen.R_SUN, 1e-6 * gen.R_SUN], (n, 1))
    zeros = np.zeros((n, 2))
    spec = gen.build_spec(np.full(n, 2), mass, radius, np.full((n, 2), 0.3), pos, vel, np.zeros((n, 2, 3)),
                          np.zeros((n, 2), dtype=np.uint32), zeros, zeros, zeros, zeros, dt,
                          abi.COORD_DEMOCRATIC_HELIOCENTRIC, effects=0)
    return spec


@pytest.mark.parametrize("mapping,group_size", [("packed", -1), ("plain", 0)])
def test_kepler_paths_bitwise_against_oracle(oracle, mapping, group_size):
    from posidonius_amd.ensemble import Ensemble
    spec = two_body_ensemble(192, 2025)
    steps = 3
    log(f"kepler paths: {spec.n_systems} two-body systems x {steps} steps, {mapping}")
    ref = oracle.run(spec, steps, threads=8)
    st = ref["status"]
    assert np.count_nonzero(st & abi.STATUS_KEPLER_BISECTION) >= 20
    assert np.count_nonzero(st & abi.STATUS_TIMESTEP_WARNING) >= 20
    with Ensemble(spec, mapping=mapping, group_size=group_size, steps_per_launch=steps) as ens:
        ens.step(steps)
        out = ens.get_state()
    keys = ("position", "velocity", "acceleration", "velocity_error", "current_time", "current_iteration",
            "status", "timestep_warning")
    assert_same_state(out, ref, keys=keys, where=f"kepler paths ({mapping})")
