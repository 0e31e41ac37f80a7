"""Seeded synthetic ensembles for the five benchmark configurations.

Each generator draws orbital elements and physical parameters with numpy's
PCG64 (seed = 1000 + configuration index unless given), converts elements
to Cartesian coordinates relative to the host (the textbook two-body
transformation of Mercury's ``mco_el2x``, the same convention as the
reference's ``calculate_cartesian_coordinates``, src/tools.rs:289-386:
``p`` is the longitude of perihelion), shifts to the barycentre the way
``Universe::new`` does (src/particles/universe.rs:97-105, 663-697) and
returns an :class:`EnsembleSpec` in the state the reference would write to
case.json before the first step.

Physical defaults follow the reference's examples: stellar CTL dissipation
2.006 * 3.845764e4 (tests/common/stars.rs, cases/example.py:43-47), planetary
sigma = 2 K2 k2dt / (3 R^5) with k2dt = 2.465278e-3 (terrestrial) or
2.893519e-7 (gas giant) (tests/common/planets.rs:101-102, 184-185),
I = m rg^2 R^2 and L = I Omega (src/particles/particle.rs:56-97).
"""
from __future__ import annotations

import numpy as np

from . import abi
from .case import EnsembleSpec

# Units and constants (src/constants.rs)
HOUR = 3600.0
DAY = 24.0 * HOUR
M_SUN = 1.9818e30
G_SI = 6.67428e-11
AU = 1.49597870700e11
K2 = G_SI / (AU * AU * AU) * M_SUN * (DAY * DAY)
G = K2
M2EARTH = 332946.050895
M_EARTH = 1.0 / M2EARTH
R_SUN = 6.957e8 / AU
R_EARTH = 6.3781e6 / AU
R_JUPITER = 7.1492e7 / AU
M_JUPITER = 9.5459e-4
DEG2RAD = np.pi / 180.0
TWO_PI = 2.0 * np.pi

STAR_DISSIPATION = 2.006 * 3.845764e4
K2DT_TERRESTRIAL = 2.465278e-3
K2DT_GAS = 2.893519e-7

ROLES_CENTRAL = abi.ROLE_TIDES_CENTRAL | abi.ROLE_FLAT_CENTRAL | abi.ROLE_GR_CENTRAL
ROLES_ORBITING = abi.ROLE_TIDES_ORBITING | abi.ROLE_FLAT_ORBITING | abi.ROLE_GR_ORBITING
FULL_EFFECTS = abi.EFFECT_TIDES | abi.EFFECT_ROTATIONAL_FLATTENING | abi.EFFECT_GENERAL_RELATIVITY


# osculating elements kept per body by planetary_system (host rows zero)
ELEMENT_KEYS = ("a", "e", "inc", "varpi", "node", "mean_anomaly")


def planet_dissipation(radius, k2dt):
    return 2.0 * K2 * k2dt / (3.0 * radius ** 5)


def solve_kepler(e, mean_anomaly):
    """Eccentric anomaly of an ellipse by Newton iteration (vectorised)."""
    M = np.mod(mean_anomaly, TWO_PI)
    E = np.where(e < 0.8, M, np.pi * np.ones_like(M))
    for _ in range(60):
        f = E - e * np.sin(E) - M
        dE = f / (1.0 - e * np.cos(E))
        E = E - dE
        if np.all(np.abs(dE) <= 4e-16 * np.maximum(1.0, np.abs(E))):
            break
    return E


def cartesian_from_elements(gm, q, e, inc, p, node, mean_anomaly):
    """Position and velocity relative to the central mass (elliptic orbits).

    p is the longitude of perihelion (argument of perihelion + node), as in
    the reference's Mercury-style interface."""
    gm, q, e, inc, p, node, mean_anomaly = np.broadcast_arrays(
        *[np.asarray(v, dtype=np.float64) for v in (gm, q, e, inc, p, node, mean_anomaly)])
    g = p - node
    si, ci = np.sin(inc), np.cos(inc)
    sg, cg = np.sin(g), np.cos(g)
    sn, cn = np.sin(node), np.cos(node)
    z1, z2, z3, z4 = cg * cn, cg * sn, sg * cn, sg * sn
    d11, d12, d13 = z1 - z4 * ci, z2 + z3 * ci, sg * si
    d21, d22, d23 = -z3 - z2 * ci, -z4 + z1 * ci, cg * si
    a = q / (1.0 - e)
    E = solve_kepler(e, mean_anomaly)
    se, ce = np.sin(E), np.cos(E)
    romes = np.sqrt(1.0 - e * e)
    x1 = a * (ce - e)
    x2 = a * romes * se
    temp = np.sqrt(gm / a) / (1.0 - e * ce)
    v1 = -se * temp
    v2 = romes * ce * temp
    pos = np.stack([d11 * x1 + d21 * x2, d12 * x1 + d22 * x2, d13 * x1 + d23 * x2], axis=-1)
    vel = np.stack([d11 * v1 + d21 * v2, d12 * v1 + d22 * v2, d13 * v1 + d23 * v2], axis=-1)
    return pos, vel


def spin_vector(angular_frequency, inclination, obliquity, node):
    """Spin vector tilted by inclination + obliquity (src/tools.rs:919-927)."""
    s = np.sin(obliquity + inclination)
    return np.stack([angular_frequency * s * np.sin(node), -angular_frequency * s * np.cos(node),
                     angular_frequency * np.cos(obliquity + inclination)], axis=-1)


def _barycentric(pos, vel, mass, counts):
    """Subtract the sequentially accumulated centre of mass (universe.rs:663-697)."""
    ns, nmax = mass.shape
    cpos = np.zeros((ns, 3))
    cvel = np.zeros((ns, 3))
    cm = np.zeros(ns)
    for j in range(nmax):
        live = j < counts
        m = mass[:, j]
        npos = cpos * cm[:, None] + pos[:, j] * m[:, None]
        nvel = cvel * cm[:, None] + vel[:, j] * m[:, None]
        new = cm + m
        safe = np.where(new > 0.0, new, 1.0)[:, None]
        npos = np.where((new > 0.0)[:, None], npos / safe, npos)
        nvel = np.where((new > 0.0)[:, None], nvel / safe, nvel)
        cpos = np.where(live[:, None], npos, cpos)
        cvel = np.where(live[:, None], nvel, cvel)
        cm = np.where(live, new, cm)
    return pos - cpos[:, None, :], vel - cvel[:, None, :]


def build_spec(counts, mass, radius, rg2, pos, vel, spin, roles, sigma, sigma_scale, k2, k2f,
               time_step, coordinates, effects=FULL_EFFECTS, time_limit=1e30, names=None) -> EnsembleSpec:
    """Flatten padded [n_systems, n_max] arrays into an EnsembleSpec.

    pos/vel are host-relative ("heliocentric" in the reference's input sense);
    they are shifted to the barycentre here."""
    counts = np.asarray(counts, dtype=np.int32)
    ns, nmax = mass.shape
    bpos, bvel = _barycentric(pos, vel, mass, counts)
    live = np.arange(nmax)[None, :] < counts[:, None]
    inertia = mass * rg2 * radius ** 2
    L = inertia[..., None] * spin
    off = np.zeros(ns + 1, dtype=np.int32)
    off[1:] = np.cumsum(counts)
    spec = EnsembleSpec(body_offset=off)
    flat = lambda a: np.ascontiguousarray(a[live])
    spec.body = {
        "mass": flat(mass), "radius": flat(radius), "radius_of_gyration_2": flat(rg2),
        "moment_of_inertia": flat(inertia), "position": flat(bpos), "velocity": flat(bvel),
        "spin": flat(spin), "angular_momentum": flat(L), "heliocentric_velocity": None,
        "roles": flat(roles).astype(np.uint32), "tides_dissipation_factor": flat(sigma),
        "tides_dissipation_factor_scale": flat(sigma_scale), "tides_love_number": flat(k2),
        "flattening_love_number": flat(k2f), "wind_k_factor": None, "wind_rotation_saturation": None,
        "velocity_error": None, "angular_momentum_error": None, "evolution_type": None,
        "evolver_left_index": None,
    }
    spec.system = {
        "coordinates": np.full(ns, coordinates, dtype=np.int32),
        "effects": np.full(ns, effects, dtype=np.uint32),
        "gr_implementation": np.full(ns, abi.GR_KIDDER1995, dtype=np.int32),
        "time_step": np.asarray(time_step, dtype=np.float64) * np.ones(ns),
        "time_limit": np.full(ns, time_limit, dtype=np.float64),
        "current_time": np.zeros(ns), "historic_snapshot_period": np.full(ns, np.inf),
        "last_historic_snapshot_time": np.full(ns, -1.0), "current_iteration": np.zeros(ns, dtype=np.int64),
    }
    spec.names = list(names) if names is not None else [f"system-{i}" for i in range(ns)]
    return spec


def _star_radius(m):
    """Low-mass main-sequence radius, anchored at 0.117 R_sun for 0.08 M_sun."""
    return 0.117 * R_SUN * (m / 0.08) ** 0.8


def planetary_system(rng, n_systems, n_planets, star_mass, planet_mass, a1, period_ratio, ecc_max, ecc_min,
                     incl_max_deg, k2_planet, sigma_scale, dt_fraction=None, dt=None, star_radius=None,
                     star_rotation_period=3.3, planet_rotation_period=1.0, names=None):
    """Star + planets on nested orbits; arrays are per system / per planet slot."""
    n_planets = np.asarray(n_planets, dtype=np.int32)
    nmax = int(n_planets.max()) + 1
    ns = n_systems
    counts = n_planets + 1
    mass = np.zeros((ns, nmax))
    radius = np.zeros((ns, nmax))
    rg2 = np.zeros((ns, nmax))
    pos = np.zeros((ns, nmax, 3))
    vel = np.zeros((ns, nmax, 3))
    spin = np.zeros((ns, nmax, 3))
    roles = np.zeros((ns, nmax), dtype=np.uint32)
    sigma = np.zeros((ns, nmax))
    sscale = np.ones((ns, nmax))
    k2 = np.zeros((ns, nmax))
    elements = {k: np.zeros((ns, nmax)) for k in ELEMENT_KEYS}
    # star
    mass[:, 0] = star_mass
    radius[:, 0] = _star_radius(star_mass) if star_radius is None else star_radius
    rg2[:, 0] = 0.2
    spin[:, 0, 2] = TWO_PI / star_rotation_period
    roles[:, 0] = ROLES_CENTRAL
    sigma[:, 0] = STAR_DISSIPATION
    k2[:, 0] = 0.307
    # planets
    a = a1.copy()
    for j in range(1, nmax):
        live = j <= n_planets
        mp = planet_mass[:, j - 1]
        rp = R_EARTH * (mp / M_EARTH) ** 0.27
        e = rng.uniform(ecc_min, ecc_max, ns)
        inc = rng.uniform(0.0, incl_max_deg * DEG2RAD, ns)
        omega = rng.uniform(0.0, TWO_PI, ns)
        node = rng.uniform(0.0, TWO_PI, ns)
        M = rng.uniform(0.0, TWO_PI, ns)
        if j > 1:
            a = a * period_ratio[:, j - 2] ** (2.0 / 3.0)
        gm = G * (star_mass + mp)
        p, v = cartesian_from_elements(gm, a * (1.0 - e), e, inc, omega + node, node, M)
        mass[:, j] = np.where(live, mp, 0.0)
        radius[:, j] = np.where(live, rp, 0.0)
        rg2[:, j] = 0.575 ** 2
        pos[:, j] = np.where(live[:, None], p, 0.0)
        vel[:, j] = np.where(live[:, None], v, 0.0)
        spin[:, j] = spin_vector(TWO_PI / planet_rotation_period, inc, 0.0, node)
        roles[:, j] = ROLES_ORBITING
        sigma[:, j] = planet_dissipation(rp, K2DT_TERRESTRIAL)
        sscale[:, j] = sigma_scale[:, j - 1]
        k2[:, j] = k2_planet[:, j - 1]
        for key, val in zip(ELEMENT_KEYS, (a, e, inc, omega + node, node, M)):
            elements[key][:, j] = np.where(live, val, 0.0)
    if dt is None:
        p_inner = TWO_PI * np.sqrt(a1 ** 3 / (G * (star_mass + planet_mass[:, 0])))
        dt = dt_fraction * p_inner
    spec = build_spec(counts, mass, radius, rg2, pos, vel, spin, roles, sigma, sscale, k2, k2.copy(),
                      dt, abi.COORD_DEMOCRATIC_HELIOCENTRIC, names=names)
    live = np.arange(nmax)[None, :] < counts[:, None]
    spec.elements = {k: np.ascontiguousarray(v[live]) for k, v in elements.items()}
    return spec


def config0(seed: int = 1000) -> EnsembleSpec:
    """M dwarf + two Earths, the single system of the CPU reference run."""
    rng = np.random.default_rng(seed)
    ns = 1
    return planetary_system(
        rng, ns, np.array([2]), np.array([0.08]), np.full((ns, 2), M_EARTH),
        np.array([0.018]), np.array([[(0.025 / 0.018) ** 1.5]]), 0.1, 0.01, 1.0,
        np.full((ns, 2), 0.305), np.full((ns, 2), 1e-6), dt=0.08, star_radius=np.array([0.117 * R_SUN]),
        names=["config0"])


def config1(n_systems: int = 4096, seed: int = 1001) -> EnsembleSpec:
    """Uniform ensemble of star + 2 planets."""
    rng = np.random.default_rng(seed)
    ns = n_systems
    star = np.exp(rng.uniform(np.log(0.08), np.log(0.6), ns))
    mp = M_EARTH * np.exp(rng.uniform(np.log(0.5), np.log(5.0), (ns, 2)))
    a1 = np.exp(rng.uniform(np.log(0.01), np.log(0.05), ns))
    ratio = rng.uniform(1.3, 2.0, (ns, 1))
    return planetary_system(rng, ns, np.full(ns, 2), star, mp, a1, ratio, 0.1, 0.01, 1.0,
                            np.full((ns, 2), 0.305), np.full((ns, 2), 1e-6), dt_fraction=0.05)


def config2(n_systems: int = 65536, seed: int = 1002, n_min: int = 2, n_max: int = 12) -> EnsembleSpec:
    """Uneven-size ensemble: 1 star + 1..11 planets in compact chains."""
    rng = np.random.default_rng(seed)
    ns = n_systems
    nb = rng.integers(n_min, n_max + 1, ns)
    npl = nb - 1
    P = n_max - 1
    star = np.full(ns, 0.09)
    mp = M_EARTH * rng.uniform(0.3, 1.2, (ns, P))
    a1 = rng.uniform(0.01, 0.015, ns)
    ratio = rng.uniform(1.2, 1.8, (ns, max(P - 1, 1)))
    k2p = rng.uniform(0.2, 0.4, (ns, P))
    scale = 10.0 ** rng.uniform(-1.0, 1.0, (ns, P))
    return planetary_system(rng, ns, npl, star, mp, a1, ratio, 0.02, 0.0, 1.0, k2p, scale, dt_fraction=0.05)


def config4(n_systems: int = 1024, seed: int = 1004) -> EnsembleSpec:
    """Adjoint workload: star + 3 planets drawn as the uneven-size ensemble."""
    return config2(n_systems, seed=seed, n_min=4, n_max=4)


def config3(n_systems: int = 16384, seed: int = 1003) -> EnsembleSpec:
    """Binary host (0.7 + 0.5 M_sun) with 1-3 circumbinary giant planets (Jacobi coordinates)."""
    rng = np.random.default_rng(seed)
    ns = n_systems
    npl = rng.integers(1, 4, ns)
    nmax = 5
    counts = npl + 2
    m1, m2 = 0.7, 0.5
    mass = np.zeros((ns, nmax))
    radius = np.zeros((ns, nmax))
    rg2 = np.zeros((ns, nmax))
    pos = np.zeros((ns, nmax, 3))
    vel = np.zeros((ns, nmax, 3))
    spin = np.zeros((ns, nmax, 3))
    roles = np.zeros((ns, nmax), dtype=np.uint32)
    sigma = np.zeros((ns, nmax))
    sscale = np.ones((ns, nmax))
    k2 = np.zeros((ns, nmax))
    elements = {k: np.zeros((ns, nmax)) for k in ELEMENT_KEYS}
    # stars: star 1 is the central body of every effect, star 2 orbits it
    abin = np.exp(rng.uniform(np.log(0.05), np.log(0.2), ns))
    ebin = rng.uniform(0.0, 0.2, ns)
    inc = rng.uniform(0.0, DEG2RAD, ns)
    omega, node, M = (rng.uniform(0.0, TWO_PI, ns) for _ in range(3))
    p2, v2 = cartesian_from_elements(G * (m1 + m2), abin * (1 - ebin), ebin, inc, omega + node, node, M)
    mass[:, 0], mass[:, 1] = m1, m2
    radius[:, 0], radius[:, 1] = 0.65 * R_SUN, 0.48 * R_SUN
    rg2[:, 0:2] = 0.07
    pos[:, 1], vel[:, 1] = p2, v2
    spin[:, 0, 2] = TWO_PI / rng.uniform(2.0, 20.0, ns)
    spin[:, 1] = spin_vector(TWO_PI / rng.uniform(2.0, 20.0, ns), inc, 0.0, node)
    roles[:, 0] = ROLES_CENTRAL
    roles[:, 1] = ROLES_ORBITING
    sigma[:, 0:2] = 4.992 * 3.845764e-2      # solar-like CTL dissipation (cases/example_binary_system.py)
    k2[:, 0:2] = 0.03
    # planets around the binary's centre of mass
    cm_pos = (m2 * p2) / (m1 + m2)
    cm_vel = (m2 * v2) / (m1 + m2)
    a = 3.5 * abin * rng.uniform(1.0, 1.5, ns)
    for j in range(2, nmax):
        live = j < counts
        mp = M_EARTH * np.exp(rng.uniform(np.log(10.0), np.log(300.0), ns))
        rp = R_JUPITER * np.minimum(1.0, (mp / M_JUPITER) ** 0.5)
        e = rng.uniform(0.0, 0.05, ns)
        incp = rng.uniform(0.0, DEG2RAD, ns)
        om, nd, Mp = (rng.uniform(0.0, TWO_PI, ns) for _ in range(3))
        if j > 2:
            a = a * rng.uniform(1.5, 3.0, ns) ** (2.0 / 3.0)
        p, v = cartesian_from_elements(G * (m1 + m2 + mp), a * (1 - e), e, incp, om + nd, nd, Mp)
        mass[:, j] = np.where(live, mp, 0.0)
        radius[:, j] = np.where(live, rp, 0.0)
        rg2[:, j] = 0.504 ** 2
        pos[:, j] = np.where(live[:, None], p + cm_pos, 0.0)
        vel[:, j] = np.where(live[:, None], v + cm_vel, 0.0)
        spin[:, j] = spin_vector(TWO_PI / 0.5, incp, 0.0, nd)
        roles[:, j] = ROLES_ORBITING
        sigma[:, j] = planet_dissipation(np.where(live, rp, 1.0), K2DT_GAS)
        k2[:, j] = 0.38
    p_bin = TWO_PI * np.sqrt(abin ** 3 / (G * (m1 + m2)))
    return build_spec(counts, mass, radius, rg2, pos, vel, spin, roles, sigma, sscale, k2, k2.copy(),
                      0.05 * p_bin, abi.COORD_JACOBI)


CONFIGS = {0: config0, 1: config1, 2: config2, 3: config3, 4: config4}
