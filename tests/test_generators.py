"""Checks of the synthetic initial-condition generators (CPU only)."""
import numpy as np
import pytest

from posidonius_amd import abi, generators as gen


def test_cartesian_known_answer():
    """src/tools.rs:933-958: a = 0.018, e = 0.1, i = 5 deg, p = n = l = 0 around
    0.08 M_sun + 1 M_earth.  The reference asserts exact values produced by its
    Nijenhuis Kepler solver; this project solves Kepler's equation by Newton,
    so the values agree to rounding level (the y, z, vx values the reference
    pins are round-off residues of size 1e-15 of the exact zeros)."""
    gm = gen.G * (gen.M_EARTH + 0.08)
    a, e = 0.018, 0.1
    pos, vel = gen.cartesian_from_elements(gm, a * (1 - e), e, 5 * gen.DEG2RAD, 0.0, 0.0, 0.0)
    x, y, z = pos
    vx, vy, vz = vel
    assert abs(x - 0.0162) <= 1e-17
    assert abs(y - 0.000000000000006571920583533768) < 1e-14
    assert abs(z - 0.0000000000000005749685486518799) < 1e-15
    assert abs(vx - -0.000000000000014818017716591765) < 1e-13
    assert vy == pytest.approx(0.03987438104619194, rel=1e-14)
    assert vz == pytest.approx(0.003488556306654768, rel=1e-14)


def test_cartesian_closed_forms():
    """Vis-viva and angular momentum of the generated state (two-body identities)."""
    rng = np.random.default_rng(0)
    n = 200
    gm = gen.G * rng.uniform(0.1, 1.0, n)
    a = rng.uniform(0.01, 1.0, n)
    e = rng.uniform(0.0, 0.9, n)
    inc = rng.uniform(0, np.pi, n)
    om, node, M = (rng.uniform(0, 2 * np.pi, n) for _ in range(3))
    pos, vel = gen.cartesian_from_elements(gm, a * (1 - e), e, inc, om + node, node, M)
    r = np.linalg.norm(pos, axis=1)
    v2 = np.sum(vel ** 2, axis=1)
    np.testing.assert_allclose(v2, gm * (2 / r - 1 / a), rtol=1e-12)
    h = np.cross(pos, vel)
    np.testing.assert_allclose(np.linalg.norm(h, axis=1), np.sqrt(gm * a * (1 - e ** 2)), rtol=1e-12)
    np.testing.assert_allclose(h[:, 2] / np.linalg.norm(h, axis=1), np.cos(inc), atol=1e-12)


@pytest.mark.parametrize("cfg,n", [(0, None), (1, 32), (2, 64), (3, 32), (4, 16)])
def test_configs_are_barycentric_and_well_formed(cfg, n):
    spec = gen.CONFIGS[cfg]() if n is None else gen.CONFIGS[cfg](n)
    off = spec.body_offset
    m = spec.body["mass"]
    for s in range(spec.n_systems):
        b0, b1 = off[s], off[s + 1]
        mm = m[b0:b1]
        com = (spec.body["position"][b0:b1] * mm[:, None]).sum(0) / mm.sum()
        mom = (spec.body["velocity"][b0:b1] * mm[:, None]).sum(0)
        scale_x = (np.abs(spec.body["position"][b0:b1]) * mm[:, None]).sum() / mm.sum()
        scale_p = (np.abs(spec.body["velocity"][b0:b1]) * mm[:, None]).sum()
        assert np.all(np.abs(com) <= 1e-14 * scale_x)
        assert np.all(np.abs(mom) <= 1e-14 * scale_p)
        assert np.argmax(mm) == 0                      # the host is body 0
        roles = spec.body["roles"][b0:b1]
        assert roles[0] & abi.ROLE_TIDES_CENTRAL and np.all(roles[1:] & abi.ROLE_TIDES_ORBITING)
    L = spec.body["angular_momentum"]
    I = spec.body["moment_of_inertia"]
    np.testing.assert_allclose(L, I[:, None] * spec.body["spin"], rtol=1e-15)


def test_config2_body_count_distribution():
    spec = gen.config2(4096)
    counts = spec.n_per_system()
    assert counts.min() == 2 and counts.max() == 12
    assert abs(counts.mean() - 7.0) < 0.2
