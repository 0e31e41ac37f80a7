"""The closed-form systems on the device: each must equal the oracle's run bit
for bit over its full length (3e3 to 1e5 steps), so the device's apsidal
advances under Kidder, Anderson and Newhall GR and under flattening, its
constant-time-lag spin relaxation and its wind braking are the ones that
tests/test_oracle_closed_forms.py checks against their closed forms.  These
single-system runs also cover effect sets the benchmark ensembles do not
have on their own: GR alone in each implementation, flattening alone, tides
alone, the wind alone (whose spin is not integrated, whfast.rs:274-275) and
the wind crossing its saturation rate, in each of the three coordinate
types."""
import numpy as np
import pytest

from posidonius_amd import abi
from tests import closed_form_cases as cases

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "angular_momentum", "spin", "velocity_error", "angular_momentum_error")


def run_both(oracle, spec, steps):
    from posidonius_amd.ensemble import Ensemble
    print(f"{steps} steps", flush=True)
    with Ensemble(spec, mapping="packed", steps_per_launch=1024, use_graph=True) as ens:
        ens.step(steps)
        out = ens.get_state()
    ref = oracle.run(spec, steps)
    for k in FIELDS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(out["status"], ref["status"])
    np.testing.assert_array_equal(out["current_time"], ref["current_time"])
    assert out["current_iteration"][0] == steps
    assert out["status"][0] & (abi.STATUS_FATAL | abi.STATUS_MIDPOINT_NOT_CONVERGED) == 0
    return out


@pytest.mark.parametrize("a,e,M", [(0.05, 0.1, 1.0), (0.03, 0.3, 1.0), (0.05, 0.1, 0.5)])
def test_kidder_case(oracle, a, e, M):
    run_both(oracle, *cases.kidder_case(a, e, M)[:2])


@pytest.mark.parametrize("implementation", ["Anderson1975", "Newhall1983"])
@pytest.mark.parametrize("a,e,M", [(0.05, 0.1, 1.0), (0.03, 0.3, 1.0), (0.05, 0.1, 0.5)])
def test_other_gr_case(oracle, implementation, a, e, M):
    run_both(oracle, *cases.other_gr_case(implementation, a, e, M)[:2])


@pytest.mark.parametrize("a,e", [(0.03, 0.2), (0.05, 0.1)])
def test_flattening_case(oracle, a, e):
    run_both(oracle, *cases.flattening_case(a, e)[:2])


@pytest.mark.parametrize("spin_over_n", [3.0, 0.2])
def test_constant_time_lag_case(oracle, spin_over_n):
    run_both(oracle, *cases.ctl_case(spin_over_n)[:2])


@pytest.mark.parametrize("effects", ["wind", "wind+tides"])
@pytest.mark.parametrize("omega0", [1.0, 3.0])
def test_wind_case(oracle, omega0, effects):
    mask = abi.EFFECT_WIND | (abi.EFFECT_TIDES if effects == "wind+tides" else 0)
    spec, steps, P = cases.wind_case(omega0, mask)
    out = run_both(oracle, spec, steps)
    if effects == "wind":   # the spin is not integrated
        np.testing.assert_array_equal(out["angular_momentum"][0], P["inertia"] * P["spin"][0, 0])


OTHER_COORDINATE_CASES = {
    "anderson": lambda c: cases.other_gr_case("Anderson1975", 0.03, 0.3, 1.0, c),
    "newhall": lambda c: cases.other_gr_case("Newhall1983", 0.03, 0.3, 1.0, c),
    "kidder": lambda c: cases.kidder_case(0.03, 0.3, 1.0, c),
    "flattening": lambda c: cases.flattening_case(0.03, 0.2, c),
    "tides_spin_down": lambda c: cases.ctl_case(3.0, c),
    "tides_spin_up": lambda c: cases.ctl_case(0.2, c),
    "wind_alone": lambda c: cases.wind_case(3.0, abi.EFFECT_WIND, c),
    "wind_across_saturation": lambda c: cases.wind_case(3.0, abi.EFFECT_WIND | abi.EFFECT_TIDES, c),
    "conservative_tide": lambda c: cases.tidal_love_case(*cases.TIDAL_LOVE_CASES[0], coordinates=c),
    "stellar_tide_decay": lambda c: cases.decay_case(1e3, c),
    "planet_flattening": lambda c: cases.planet_flattening_case(0.01, 0.1, c),
    "spin_orbit": lambda c: cases.spin_orbit_case(0.02, c),
}


@pytest.mark.parametrize("coordinates", ["Jacobi", "WHDS"])
@pytest.mark.parametrize("case", list(OTHER_COORDINATE_CASES))
def test_case_in_jacobi_and_whds_coordinates(oracle, case, coordinates):
    """The same systems in the two other coordinate types (whose closed-form
    bounds tests/test_oracle_closed_forms.py checks on the oracle)."""
    spec, steps, _ = OTHER_COORDINATE_CASES[case](abi.COORDINATES[coordinates])
    run_both(oracle, spec, steps)


def _adjoint(spec, steps, w_L):
    from posidonius_amd.ensemble import Ensemble
    print(f"adjoint over {steps} steps", flush=True)
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=1000)
        st = ens.get_state()
        g = ens.adjoint(angular_momentum=w_L)
    return st, g


@pytest.mark.parametrize("spin_over_n", [3.0, 0.2])
def test_adjoint_spin_relaxation_derivatives(oracle, spin_over_n):
    """The device's discrete adjoint on the tidal spin relaxation: the
    gradient of the final spin angular momentum along the orbit normal with
    respect to the planet's dissipation factor and its initial spin equals
    the oracle's tangent-linear derivatives to 1e-6 and the closed form to
    1e-3."""
    from tests.test_oracle_closed_forms import ctl_closed_form_derivatives
    spec, steps, P = cases.ctl_case(spin_over_n)
    nb = spec.n_bodies
    w = np.zeros((nb, 3))
    w[1] = P["normal"]
    st, g = _adjoint(spec, steps, w)
    ref = oracle.run(spec, steps)
    np.testing.assert_array_equal(st["angular_momentum"], ref["angular_momentum"])
    dirs = {"sigma": np.zeros((2, nb)), "L": np.zeros((2, nb, 3))}
    dirs["sigma"][0, 1] = 1.0
    dirs["L"][1, 1] = P["normal"]
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp_sigma = tan["L"][0, 1] @ P["normal"]
    jvp_L0 = tan["L"][1, 1] @ P["normal"]
    vjp_sigma = g["dissipation_factor"][1]
    vjp_L0 = g["angular_momentum"][1] @ P["normal"]
    d_sigma, d_L0 = ctl_closed_form_derivatives(st, P)
    print(f"d/dsigma: adjoint {vjp_sigma:.9e}, tangent {jvp_sigma:.9e}, closed form {d_sigma:.9e}; "
          f"d/dL0: adjoint {vjp_L0:.9e}, tangent {jvp_L0:.9e}, closed form {d_L0:.9e}", flush=True)
    assert abs(vjp_sigma / jvp_sigma - 1) < 1e-6
    assert abs(vjp_L0 / jvp_L0 - 1) < 1e-6
    assert abs(vjp_sigma / d_sigma - 1) < 1e-3
    assert abs(vjp_L0 / d_L0 - 1) < 1e-3


@pytest.mark.parametrize("omega0", [1.0, 3.0])
def test_adjoint_wind_derivative(oracle, omega0):
    """The device's discrete adjoint of the wind's spin-down: dOmega(t)/dOmega0
    equals the oracle's tangent-linear derivative to 1e-6 and the closed form
    to 1e-6 unsaturated, 1e-3 across the saturation rate."""
    from tests.test_oracle_closed_forms import wind_closed_form_derivative
    spec, steps, P = cases.wind_case(omega0, abi.EFFECT_WIND | abi.EFFECT_TIDES)
    nb = spec.n_bodies
    w = np.zeros((nb, 3))
    w[0] = P["axis"]
    st, g = _adjoint(spec, steps, w)
    dirs = {"L": np.zeros((1, nb, 3))}
    dirs["L"][0, 0] = P["axis"]
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp = tan["L"][0, 0] @ P["axis"]
    vjp = g["angular_momentum"][0] @ P["axis"]
    closed = wind_closed_form_derivative(omega0, P, st["current_time"][0])
    print(f"dOmega/dOmega0: adjoint {vjp:.12e}, tangent {jvp:.12e}, closed form {closed:.12e}", flush=True)
    assert abs(vjp / jvp - 1) < 1e-6
    assert abs(vjp / closed - 1) < (1e-6 if omega0 < P["wsat"] else 1e-3)


@pytest.mark.parametrize("a,e", [(0.03, 0.2), (0.05, 0.1)])
def test_adjoint_flattening_advance_derivative(oracle, a, e):
    """The device's discrete adjoint, seeded with the gradient of the apsidal
    advance with respect to the final state: its gradient with respect to the
    star's fluid Love number equals the oracle's tangent-linear derivative to
    1e-6 and the closed form, the advance over k2f, to 2e-5."""
    from tests.test_oracle_closed_forms import advance_gradient
    spec, steps, P = cases.flattening_case(a, e)
    nb = spec.n_bodies
    ref = oracle.run(spec, steps)
    A, gr, gv = advance_gradient(spec, ref, P["mu"])
    from posidonius_amd.ensemble import Ensemble
    print(f"adjoint over {steps} steps", flush=True)
    wx, wv = np.zeros((nb, 3)), np.zeros((nb, 3))
    wx[1], wx[0], wv[1], wv[0] = gr, -gr, gv, -gv
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=1000)
        st = ens.get_state()
        g = ens.adjoint(position=wx, velocity=wv)
    np.testing.assert_array_equal(st["position"], ref["position"])
    dirs = {"k2f": np.zeros((1, nb))}
    dirs["k2f"][0, 0] = 1.0
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp = gr @ (tan["x"][0, 1] - tan["x"][0, 0]) + gv @ (tan["v"][0, 1] - tan["v"][0, 0])
    vjp = g["fluid_love_number"][0]
    print(f"dA/dk2f: adjoint {vjp:.12e}, tangent {jvp:.12e}, A/k2f {A / P['k2f']:.12e}", flush=True)
    assert abs(vjp / jvp - 1) < 1e-6
    assert abs(vjp / (A / P["k2f"]) - 1) < 2e-5


@pytest.mark.parametrize("case", [0, 1])
def test_conservative_tide_case_and_adjoint(oracle, case):
    """The conservative-tide system on the device, bitwise against the oracle,
    and the device adjoint's gradient of its apsidal advance with respect to
    the planet's Love number against the oracle's tangent-linear derivative
    (1e-6); tests/test_oracle_closed_forms.py checks both against the closed
    form."""
    from tests.test_oracle_closed_forms import advance_gradient
    spec, steps, P = cases.tidal_love_case(*cases.TIDAL_LOVE_CASES[case])
    out = run_both(oracle, spec, steps)
    nb = spec.n_bodies
    A, gr, gv = advance_gradient(spec, out, P["mu"])
    wx, wv = np.zeros((nb, 3)), np.zeros((nb, 3))
    wx[1], wx[0], wv[1], wv[0] = gr, -gr, gv, -gv
    from posidonius_amd.ensemble import Ensemble
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=1000)
        g = ens.adjoint(position=wx, velocity=wv)
    dirs = {"k2": np.zeros((1, nb))}
    dirs["k2"][0, 1] = 1.0
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp = gr @ (tan["x"][0, 1] - tan["x"][0, 0]) + gv @ (tan["v"][0, 1] - tan["v"][0, 0])
    vjp = g["love_number"][1]
    print(f"dA/dk2: adjoint {vjp:.12e}, tangent {jvp:.12e}, A/k2 {A / P['k2']:.12e}", flush=True)
    assert abs(vjp / jvp - 1) < 1e-6


@pytest.mark.parametrize("k2dt", [1e3, 3e3])
def test_stellar_tide_decay_case_and_adjoint(oracle, k2dt):
    """The orbital-decay system on the device, bitwise against the oracle, and
    the device adjoint's gradient of the final semi-major axis with respect
    to the star's dissipation factor against the oracle's tangent-linear
    derivative (1e-6); tests/test_oracle_closed_forms.py checks both against
    the closed form."""
    from tests.test_oracle_closed_forms import semi_major_axis_and_gradient
    spec, steps, P = cases.decay_case(k2dt)
    out = run_both(oracle, spec, steps)
    nb = spec.n_bodies
    _, ga_r, ga_v = semi_major_axis_and_gradient(out, P["mu"])
    wx, wv = np.zeros((nb, 3)), np.zeros((nb, 3))
    wx[1], wx[0], wv[1], wv[0] = ga_r, -ga_r, ga_v, -ga_v
    from posidonius_amd.ensemble import Ensemble
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=1000)
        g = ens.adjoint(position=wx, velocity=wv)
    dirs = {"sigma": np.zeros((1, nb))}
    dirs["sigma"][0, 0] = 1.0
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp = ga_r @ (tan["x"][0, 1] - tan["x"][0, 0]) + ga_v @ (tan["v"][0, 1] - tan["v"][0, 0])
    vjp = g["dissipation_factor"][0]
    print(f"da/dsigma_star: adjoint {vjp:.12e}, tangent {jvp:.12e}", flush=True)
    assert abs(vjp / jvp - 1) < 1e-6


@pytest.mark.parametrize("a,e", [(0.01, 0.1), (0.015, 0.3)])
def test_planet_flattening_case_and_adjoint(oracle, a, e):
    """The planet's-bulge system on the device, bitwise against the oracle,
    and the device adjoint's gradient of the apsidal advance with respect to
    the planet's fluid Love number against the oracle's tangent-linear
    derivative (1e-6); tests/test_oracle_closed_forms.py checks both against
    the closed form."""
    from tests.test_oracle_closed_forms import advance_gradient
    spec, steps, P = cases.planet_flattening_case(a, e)
    out = run_both(oracle, spec, steps)
    nb = spec.n_bodies
    _, gr, gv = advance_gradient(spec, out, P["mu"])
    wx, wv = np.zeros((nb, 3)), np.zeros((nb, 3))
    wx[1], wx[0], wv[1], wv[0] = gr, -gr, gv, -gv
    from posidonius_amd.ensemble import Ensemble
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=1000)
        g = ens.adjoint(position=wx, velocity=wv)
    dirs = {"k2f": np.zeros((1, nb))}
    dirs["k2f"][0, 1] = 1.0
    _, tan = oracle.tangent(spec, steps, dirs)
    jvp = gr @ (tan["x"][0, 1] - tan["x"][0, 0]) + gv @ (tan["v"][0, 1] - tan["v"][0, 0])
    vjp = g["fluid_love_number"][1]
    print(f"dA/dk2f_planet: adjoint {vjp:.12e}, tangent {jvp:.12e}", flush=True)
    assert abs(vjp / jvp - 1) < 1e-6


@pytest.mark.parametrize("a", [0.02, 0.05])
def test_spin_orbit_case(oracle, a):
    """The spin-orbit system (a star spinning about an axis tilted from the
    orbit normal, Kidder GR alone) on the device, bitwise against the
    oracle, whose nodal precession tests/test_oracle_closed_forms.py checks."""
    run_both(oracle, *cases.spin_orbit_case(a)[:2])
