"""Star + planet systems whose orbit or spin evolution has a closed form,
one builder per effect.  Each returns (spec, steps, params): the ensemble of
one system, the number of steps to run it and the quantities its closed form
needs.  `coordinates` is the integrator's coordinate type (democratic
heliocentric by default).  tests/test_oracle_closed_forms.py checks the oracle's runs of them
against the closed forms and independent integrations; the device's runs
must equal the oracle's bit for bit (tests/test_gpu_closed_forms.py)."""
import numpy as np

from posidonius_amd import abi, generators as gen


def _orbit(mu, a, e, varpi, node, mean_anomaly):
    """Host-relative planet orbit in the xy plane (inclination 0), varpi the
    longitude of perihelion."""
    p, v = gen.cartesian_from_elements(mu, np.array([a * (1 - e)]), np.array([e]), np.array([0.0]),
                                       np.array([varpi]), np.array([node]), np.array([mean_anomaly]))
    pos = np.zeros((1, 2, 3)); vel = np.zeros((1, 2, 3))
    pos[0, 1], vel[0, 1] = p[0], v[0]
    return pos, vel


def kidder_case(a, e, M, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only Kidder GR on; 300 orbits of 100 steps.  The spins lie along z,
    slow enough that their spin-orbit terms are below 1e-8 of the advance
    (with every spin zero the midpoint's convergence test divides zero by
    zero, as the reference's does)."""
    m = gen.M_EARTH
    mu = gen.G * (M + m)
    period = 2 * np.pi * np.sqrt(a ** 3 / mu)
    orbits, steps_per_orbit = 300, 100
    pos, vel = _orbit(mu, a, e, 0.7, 0.0, 0.4)
    zeros = np.zeros((1, 2))
    roles = np.array([[abi.ROLE_GR_CENTRAL, abi.ROLE_GR_ORBITING]], dtype=np.uint32)
    spin = np.zeros((1, 2, 3))
    spin[0, :, 2] = 2 * np.pi / 1e6
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[gen.R_SUN, gen.R_EARTH]]),
                          np.array([[0.07, 0.33]]), pos, vel, spin, roles, zeros, np.ones((1, 2)),
                          zeros, zeros, period / steps_per_orbit, coordinates,
                          effects=abi.EFFECT_GENERAL_RELATIVITY)
    return spec, orbits * steps_per_orbit, dict(M=M, m=m, mu=mu, period=period)


def other_gr_case(implementation, a, e, M, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only GR on, with the Anderson 1975 or Newhall 1983 implementation (no
    spin integration); 300 orbits of 100 steps."""
    m = gen.M_EARTH
    mu = gen.G * (M + m)
    period = 2 * np.pi * np.sqrt(a ** 3 / mu)
    orbits, steps_per_orbit = 300, 100
    pos, vel = _orbit(mu, a, e, 0.7, 0.0, 0.4)
    zeros = np.zeros((1, 2))
    roles = np.array([[abi.ROLE_GR_CENTRAL, abi.ROLE_GR_ORBITING]], dtype=np.uint32)
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[gen.R_SUN, gen.R_EARTH]]),
                          np.array([[0.07, 0.33]]), pos, vel, np.zeros((1, 2, 3)), roles, zeros, np.ones((1, 2)),
                          zeros, zeros, period / steps_per_orbit, coordinates,
                          effects=abi.EFFECT_GENERAL_RELATIVITY)
    spec.system["gr_implementation"][:] = abi.GR_IMPLEMENTATIONS[implementation]
    return spec, orbits * steps_per_orbit, dict(M=M, m=m, mu=mu, period=period)


def flattening_case(a, e, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only rotational flattening on, the star spinning along the normal of
    an orbit in its equatorial plane (the z axis), the planet's fluid Love
    number zero;
    1,000 orbits of 100 steps."""
    M, m = 1.0, gen.M_EARTH
    R, k2f, omega = gen.R_SUN, 0.03, 2 * np.pi / 2.0
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    orbits, steps_per_orbit = 1000, 100
    pos, vel = _orbit(mu, a, e, 1.1, 0.0, 2.0)
    roles = np.array([[abi.ROLE_FLAT_CENTRAL, abi.ROLE_FLAT_ORBITING]], dtype=np.uint32)
    normal = np.cross(pos[0, 1], vel[0, 1])
    normal /= np.linalg.norm(normal)
    spin = np.zeros((1, 2, 3))
    spin[0, 0] = omega * normal
    spin[0, 1] = 2 * np.pi / 1.0 * normal
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[R, gen.R_EARTH]]), np.array([[0.07, 0.33]]),
                          pos, vel, spin, roles, zeros, np.ones((1, 2)), zeros, np.array([[k2f, 0.0]]),
                          period / steps_per_orbit, coordinates,
                          effects=abi.EFFECT_ROTATIONAL_FLATTENING)
    return spec, orbits * steps_per_orbit, dict(M=M, m=m, mu=mu, n=n, period=period, R=R, k2f=k2f, omega=omega)


def ctl_case(spin_over_n, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only tides on, dissipating in the planet alone (sigma = 2 G k2 dt /
    (3 R^5), as the reference's example cases set it), a circular orbit and
    the spins along its normal; 1.5 spin-relaxation times at 50 steps per
    orbit."""
    M, m, R, rg2, a = 0.08, gen.M_EARTH, gen.R_EARTH, 0.33, 0.01
    k2dt = 10 * 2.465278e-3
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    tau = rg2 * m * a ** 6 / (3 * gen.G * M ** 2 * k2dt * R ** 3)
    steps_per_orbit = 50
    steps = int(round(1.5 * tau / (period / steps_per_orbit)))
    pos, vel = _orbit(mu, a, 0.0, 0.6, 0.0, 1.3)
    normal = np.cross(pos[0, 1], vel[0, 1])
    normal /= np.linalg.norm(normal)
    spin = np.zeros((1, 2, 3))
    spin[0, 0] = 2 * np.pi / 30.0 * normal
    spin[0, 1] = spin_over_n * n * normal
    roles = np.array([[abi.ROLE_TIDES_CENTRAL, abi.ROLE_TIDES_ORBITING]], dtype=np.uint32)
    sigma = np.array([[0.0, 2 * gen.G * k2dt / (3 * R ** 5)]])
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[0.1 * gen.R_SUN, R]]), np.array([[0.2, rg2]]),
                          pos, vel, spin, roles, sigma, np.ones((1, 2)), zeros, zeros,
                          period / steps_per_orbit, coordinates, effects=abi.EFFECT_TIDES)
    return spec, steps, dict(M=M, m=m, R=R, rg2=rg2, a=a, k2dt=k2dt, mu=mu, n=n, period=period, tau=tau,
                             normal=normal, spin=spin)


def wind_case(omega0, effects, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """A 0.8 solar-mass star of 0.9 solar radii braked by its wind (k giving a
    spin-down time of order 1e3 days at 1 rad/day, saturation at 2 rad/day),
    with a planet at 0.1 AU; 3,000 days at 20 steps per orbit.  effects is
    the wind alone or the wind with tides of zero dissipation and Love
    numbers (no tidal force or torque)."""
    M, R, rg2, m, a, wsat = 0.8, 0.9 * gen.R_SUN, 0.07, gen.M_EARTH, 0.1, 2.0
    mu = gen.G * (M + m)
    period = 2 * np.pi * np.sqrt(a ** 3 / mu)
    inertia = rg2 * M * R ** 2
    k = inertia / 2e3
    pos, vel = _orbit(mu, a, 0.0, 0.3, 0.0, 0.4)
    axis = np.array([0.6, 0.0, 0.8])
    spin = np.zeros((1, 2, 3))
    spin[0, 0] = omega0 * axis
    spin[0, 1, 2] = 1.0
    roles = np.array([[abi.ROLE_WIND | abi.ROLE_TIDES_CENTRAL, abi.ROLE_TIDES_ORBITING]], dtype=np.uint32)
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[R, gen.R_EARTH]]), np.array([[rg2, 0.33]]),
                          pos, vel, spin, roles, zeros, np.ones((1, 2)), zeros, zeros, period / 20,
                          coordinates, effects=effects)
    spec.body["wind_k_factor"] = np.array([k, 0.0])
    spec.body["wind_rotation_saturation"] = np.array([wsat, 0.0])
    steps = int(round(3e3 / (period / 20)))
    return spec, steps, dict(M=M, R=R, m=m, inertia=inertia, k=k, wsat=wsat, axis=axis, spin=spin,
                             rate=k * np.sqrt(R / gen.R_SUN / M) / inertia)


# (M, m, R, a, e) of the conservative-tide systems: an Earth around an M
# dwarf, a Neptune-mass planet on an eccentric orbit around a Sun
TIDAL_LOVE_CASES = [(0.08, gen.M_EARTH, gen.R_EARTH, 0.01, 0.1), (1.0, 1e-4, 0.05 * gen.R_SUN, 0.04, 0.3)]


def tidal_love_case(M, m, R, a, e, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only tides on with zero dissipation: the conservative tide the star
    raises on the planet (Love number 0.3; the star's is zero), the spins
    along the orbit normal (no torque without dissipation); 1,000 orbits of
    100 steps."""
    k2 = 0.3
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    orbits, steps_per_orbit = 1000, 100
    pos, vel = _orbit(mu, a, e, 0.7, 0.0, 0.4)
    spin = np.zeros((1, 2, 3))
    spin[0, 0, 2] = 2 * np.pi / 30.0
    spin[0, 1, 2] = n
    roles = np.array([[abi.ROLE_TIDES_CENTRAL, abi.ROLE_TIDES_ORBITING]], dtype=np.uint32)
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[gen.R_SUN, R]]), np.array([[0.07, 0.25]]),
                          pos, vel, spin, roles, zeros, np.ones((1, 2)), np.array([[0.0, k2]]), zeros,
                          period / steps_per_orbit, coordinates, effects=abi.EFFECT_TIDES)
    return spec, orbits * steps_per_orbit, dict(M=M, m=m, R=R, a=a, e=e, k2=k2, mu=mu, n=n, period=period)


def decay_case(k2dt, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only tides on, dissipating in the star alone (sigma = 2 G k2 dt /
    (3 R^5) with the given k2 dt, days; the Love numbers zero), an Earth-mass
    planet on a circular orbit at 0.02 AU around a solar-mass star that
    barely rotates (1e-6 of the mean motion, along the orbit normal);
    5,000 steps of a hundredth of the period."""
    M, m, Rs, a = 1.0, gen.M_EARTH, gen.R_SUN, 0.02
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    pos, vel = _orbit(mu, a, 0.0, 0.3, 0.0, 0.9)
    spin = np.zeros((1, 2, 3))
    spin[0, 0, 2] = 1e-6 * n
    spin[0, 1, 2] = n
    roles = np.array([[abi.ROLE_TIDES_CENTRAL, abi.ROLE_TIDES_ORBITING]], dtype=np.uint32)
    zeros = np.zeros((1, 2))
    sigma = 2 * gen.G * k2dt / (3 * Rs ** 5)
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[Rs, gen.R_EARTH]]), np.array([[0.07, 0.33]]),
                          pos, vel, spin, roles, np.array([[sigma, 0.0]]), np.ones((1, 2)), zeros, zeros,
                          period / 100, coordinates, effects=abi.EFFECT_TIDES)
    return spec, 5000, dict(M=M, m=m, Rs=Rs, a=a, k2dt=k2dt, sigma=sigma, mu=mu, n=n,
                            inertia=0.07 * M * Rs ** 2)


def planet_flattening_case(a, e, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only rotational flattening on, the planet's own bulge: an Earth-like
    planet (fluid Love number 0.3, one-day spin along the normal of its
    orbit, the z axis) around a 0.08 solar-mass star whose fluid Love number
    is zero; 1,000 orbits of 100 steps."""
    M, m, R = 0.08, gen.M_EARTH, gen.R_EARTH
    k2f, omega = 0.3, 2 * np.pi / 1.0
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    pos, vel = _orbit(mu, a, e, 1.1, 0.0, 2.0)
    roles = np.array([[abi.ROLE_FLAT_CENTRAL, abi.ROLE_FLAT_ORBITING]], dtype=np.uint32)
    spin = np.zeros((1, 2, 3))
    spin[0, 0, 2] = 2 * np.pi / 30.0
    spin[0, 1, 2] = omega
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[0.1 * gen.R_SUN, R]]), np.array([[0.2, 0.33]]),
                          pos, vel, spin, roles, zeros, np.ones((1, 2)), zeros, np.array([[0.0, k2f]]),
                          period / 100, coordinates, effects=abi.EFFECT_ROTATIONAL_FLATTENING)
    return spec, 100 * 1000, dict(M=M, m=m, R=R, k2f=k2f, omega=omega, mu=mu, n=n, period=period)


def spin_orbit_case(a, coordinates=abi.COORD_DEMOCRATIC_HELIOCENTRIC):
    """Only Kidder GR on, an Earth-mass planet on a circular orbit around a
    solar-mass star spinning with a 10-day period about an axis 30 degrees
    from the orbit normal (the z axis); the planet's spin slow (1e-6 rad/day)
    along z; 1e5 steps of a hundredth of the period."""
    M, m = 1.0, gen.M_EARTH
    mu = gen.G * (M + m)
    n = np.sqrt(mu / a ** 3)
    period = 2 * np.pi / n
    pos, vel = _orbit(mu, a, 0.0, 0.7, 0.0, 0.4)
    tilt = np.radians(30.0)
    axis = np.array([np.sin(tilt), 0.0, np.cos(tilt)])
    spin = np.zeros((1, 2, 3))
    spin[0, 0] = 2 * np.pi / 10.0 * axis
    spin[0, 1, 2] = 2 * np.pi / 1e6
    roles = np.array([[abi.ROLE_GR_CENTRAL, abi.ROLE_GR_ORBITING]], dtype=np.uint32)
    zeros = np.zeros((1, 2))
    spec = gen.build_spec([2], np.array([[M, m]]), np.array([[gen.R_SUN, gen.R_EARTH]]), np.array([[0.07, 0.33]]),
                          pos, vel, spin, roles, zeros, np.ones((1, 2)), zeros, zeros, period / 100, coordinates,
                          effects=abi.EFFECT_GENERAL_RELATIVITY)
    return spec, 100000, dict(M=M, m=m, a=a, mu=mu, n=n, period=period, axis=axis)
