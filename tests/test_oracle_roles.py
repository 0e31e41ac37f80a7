"""Per-body effect roles in the oracle, pinned by what the reference's effect
loops fix rather than by a fixture (no reference fixture has a body whose
tides, flattening or general relativity is Disabled).

In the reference each effect loops over its orbiting bodies only
(constant_time_lag.rs:217-218, oblate_spheroid.rs:24, general_relativity.rs:
96-97) and the host takes the reaction of those bodies.  So, with one effect
on and planet k's role for it cleared, one evaluation of the additional
effects must give:

* planet k: exactly the acceleration and spin derivative it gets with the
  effect switched off;
* every other planet: exactly what it gets with all roles set;
* the host: its all-roles value less the reaction to planet k alone, which
  is the host's value in the two-body system (host, planet k), to rounding;
* for tides only, the host's spin derivative is NaN: the host's torque loop
  tests the host's effect, not the body's (tides/common.rs:244-255), and the
  skipped body's tidal distance, position and components are the zeros the
  case was written with, so its term is 0 * (0 - 0 / 0 - 1 / 0 * 0).
"""
import numpy as np
import pytest

from posidonius_amd import abi
from posidonius_amd.case import EnsembleSpec
from tests.oracle_runner import load_golden

FIXTURE = "test_tides-enabled_tides"   # star + 4 planets, tides + flattening + Kidder GR
EFFECTS = {"tides": (abi.EFFECT_TIDES, abi.ROLE_TIDES_ORBITING),
           "flattening": (abi.EFFECT_ROTATIONAL_FLATTENING, abi.ROLE_FLAT_ORBITING),
           "general_relativity": (abi.EFFECT_GENERAL_RELATIVITY, abi.ROLE_GR_ORBITING)}
ALL = abi.EFFECT_TIDES | abi.EFFECT_ROTATIONAL_FLATTENING | abi.EFFECT_GENERAL_RELATIVITY


def base(effects):
    """The fixture at t = 0 with the given effects and no evolution (so that
    the radii are the case's own in every variant)."""
    spec, _ = load_golden(FIXTURE)
    spec.system["effects"] = np.array([effects], dtype=np.uint32)
    spec.body["evolution_type"][:] = abi.EVO_NON_EVOLVING
    spec.evolver_offset = None
    spec.tables = {}
    return spec


def without_role(spec, body, role):
    spec = spec.copy()
    spec.body["roles"][body] &= ~np.uint32(role)
    return spec


def pair_system(spec, host, k):
    """The two-body system (host, body k) with the same inertial state."""
    keep = np.array([host, k])
    out = EnsembleSpec(body_offset=np.array([0, 2], dtype=np.int32))
    out.system = {key: (None if v is None else v.copy()) for key, v in spec.system.items()}
    out.body = {key: (None if v is None else v[keep].copy()) for key, v in spec.body.items()}
    return out


@pytest.mark.parametrize("effect", list(EFFECTS))
@pytest.mark.parametrize("k", [1, 3])
def test_disabled_role_alone(oracle, effect, k):
    bit, role = EFFECTS[effect]
    spec = base(bit)
    a_all, l_all = oracle.additional_effects(spec)
    a_off, l_off = oracle.additional_effects(base(0))
    a_dis, l_dis = oracle.additional_effects(without_role(spec, k, role))
    a_pair, l_pair = oracle.additional_effects(pair_system(spec, 0, k))
    # planet k: its effect-off values, bit for bit
    np.testing.assert_array_equal(a_dis[k], a_off[k])
    np.testing.assert_array_equal(l_dis[k], l_off[k])
    # the other planets: untouched
    for j in range(1, spec.n_bodies):
        if j != k:
            np.testing.assert_array_equal(a_dis[j], a_all[j])
            np.testing.assert_array_equal(l_dis[j], l_all[j])
    # the pair system reproduces planet k's own all-roles terms
    np.testing.assert_array_equal(a_pair[1], a_all[k])
    # the host's acceleration: all-roles value less planet k's reaction, to
    # the rounding of sums over four bodies (a few ulps of the host's total)
    react = a_pair[0]
    atol = 16 * np.finfo(float).eps * np.abs(a_all[0]).max()
    assert np.abs(react).max() > 1e4 * atol, "the reaction to planet k must be resolvable"
    np.testing.assert_allclose(a_all[0] - a_dis[0], react, rtol=0, atol=atol)
    if effect == "tides":
        assert np.all(np.isnan(l_dis[0])), "the reference's host tidal torque includes a NaN term"
    else:
        atol = 16 * np.finfo(float).eps * np.abs(l_all[0]).max()
        assert np.abs(l_pair[0]).max() > 1e4 * atol
        np.testing.assert_allclose(l_all[0] - l_dis[0], l_pair[0], rtol=0, atol=atol)


def test_all_effects_with_one_role_cleared_each(oracle):
    """Tides, flattening and GR on together; planet 1 without tides, planet 2
    without flattening, planet 3 without GR: each planet's other effects are
    those of the all-roles evaluation to rounding of the sums it is part of."""
    spec = base(ALL)
    a_all, l_all = oracle.additional_effects(spec)
    s = without_role(spec, 1, abi.ROLE_TIDES_ORBITING)
    s = without_role(s, 2, abi.ROLE_FLAT_ORBITING)
    s = without_role(s, 3, abi.ROLE_GR_ORBITING)
    a, l = oracle.additional_effects(s)
    np.testing.assert_array_equal(a[4], a_all[4])       # planet 4 keeps every role
    np.testing.assert_array_equal(l[4], l_all[4])
    assert np.all(np.isnan(l[0]))                       # tides on with a skipped body
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(l[1:]))
    # planet 1 equals the evaluation with tides off for it: flattening + GR only
    a_rg, l_rg = oracle.additional_effects(base(abi.EFFECT_ROTATIONAL_FLATTENING | abi.EFFECT_GENERAL_RELATIVITY))
    np.testing.assert_array_equal(a[1], a_rg[1])
    np.testing.assert_array_equal(l[1], l_rg[1])


def test_effect_without_central_body_is_switched_off(oracle):
    """disable_unnecessary_effects (universe.rs:699-764): tides on but no
    tidal central body evaluates as tides off."""
    spec = base(ALL)
    no_host = without_role(spec, 0, abi.ROLE_TIDES_CENTRAL)
    a, l = oracle.additional_effects(no_host)
    a_off, l_off = oracle.additional_effects(base(abi.EFFECT_ROTATIONAL_FLATTENING | abi.EFFECT_GENERAL_RELATIVITY))
    np.testing.assert_array_equal(a, a_off)
    np.testing.assert_array_equal(l, l_off)


def test_disabled_role_runs_steps(oracle):
    """A few steps with a flattening- and GR-disabled planet stay finite and
    differ from the all-roles run; with a tides-disabled planet the host's
    spin turns NaN in the first kick, as the reference's does."""
    spec = base(ALL)
    s = without_role(without_role(spec, 2, abi.ROLE_FLAT_ORBITING), 3, abi.ROLE_GR_ORBITING)
    out = oracle.run(s, 20)
    ref = oracle.run(spec, 20)
    assert np.all(np.isfinite(out["position"]))
    assert not np.array_equal(out["position"], ref["position"])
    bad = oracle.run(without_role(spec, 2, abi.ROLE_TIDES_ORBITING), 1)
    assert np.all(np.isnan(bad["angular_momentum"][0]))
