"""Closed-form two-body propagation, independent of the universal-variable
formulation of the drift under test: the osculating elements of (x, v), the
mean anomaly advanced by n dt, Kepler's equation solved for the eccentric
(E - e sin E = M) or hyperbolic (e sinh H - H = M) anomaly, and the position
and velocity rebuilt in the perifocal frame."""
import numpy as np


def two_body(gm, dt, x, v):
    x = np.asarray(x, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    r = np.linalg.norm(x)
    h = np.cross(x, v)
    evec = np.cross(v, h) / gm - x / r
    e = np.linalg.norm(evec)
    a = 1.0 / (2.0 / r - v @ v / gm)
    P = evec / e
    Q = np.cross(h, P) / np.linalg.norm(h)
    if a > 0.0:
        n = np.sqrt(gm / a ** 3)
        E0 = np.arctan2((x @ v) / np.sqrt(gm * a), 1.0 - r / a)   # (e sin E, e cos E)
        M = E0 - e * np.sin(E0) + n * dt
        Mr = np.mod(M, 2.0 * np.pi)
        E = Mr + 0.85 * e * np.sign(np.sin(Mr))                    # Danby's starting value
        for _ in range(100):
            E = E - (E - e * np.sin(E) - Mr) / (1.0 - e * np.cos(E))
        assert abs(E - e * np.sin(E) - Mr) < 1e-12 * max(1.0, abs(Mr))
        b = a * np.sqrt(1.0 - e * e)
        Edot = n / (1.0 - e * np.cos(E))
        X = a * (np.cos(E) - e) * P + b * np.sin(E) * Q
        V = -a * np.sin(E) * Edot * P + b * np.cos(E) * Edot * Q
    else:
        n = np.sqrt(gm / (-a) ** 3)
        H0 = np.arcsinh((x @ v) / np.sqrt(-gm * a) / e)               # e sinh H = x.v / sqrt(-gm a)
        M = e * np.sinh(H0) - H0 + n * dt
        H = np.sign(M) * np.log(2.0 * abs(M) / e + 1.8)
        for _ in range(200):
            H = H - (e * np.sinh(H) - H - M) / (e * np.cosh(H) - 1.0)
        assert abs(e * np.sinh(H) - H - M) < 1e-12 * max(1.0, abs(M))
        b = -a * np.sqrt(e * e - 1.0)
        Hdot = n / (e * np.cosh(H) - 1.0)
        X = a * (np.cosh(H) - e) * P + b * np.sinh(H) * Q
        V = a * np.sinh(H) * Hdot * P + b * np.cosh(H) * Hdot * Q
    return X, V, a, e


def random_orbit(rng, gm, speed_lo, speed_hi):
    """A position at 0.01-2 AU and a velocity of speed_lo..speed_hi times the
    circular speed in a random direction (hyperbolic above sqrt(2))."""
    r = rng.uniform(0.01, 2.0)
    x = rng.normal(size=3)
    x *= r / np.linalg.norm(x)
    v = rng.normal(size=3)
    v -= (v @ x) / (x @ x) * x * rng.uniform(0.0, 1.0)
    v *= np.sqrt(gm / r) * rng.uniform(speed_lo, speed_hi) / np.linalg.norm(v)
    return x, v
