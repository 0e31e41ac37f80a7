"""The discrete adjoint (pos_adjoint_record / pos_adjoint) against the
oracle's tangent-linear integration of the same steps.

For a seed w on the final state and a direction d on the initial state and
parameters, the adjoint's gradient g must satisfy g . d = w . (J d), where
J d is the tangent-linear (variational) solution of the restated step (the
oracle instantiated with dual numbers).  Directions and seeds are taken one
block at a time so that an error in any single block shows on its own.
"""
import os

import numpy as np
import pytest

from posidonius_amd import generators as gen
from tests.spec_tools import move_host

pytestmark = pytest.mark.gpu

BLOCKS = ("position", "velocity", "angular_momentum")


@pytest.fixture(scope="module")
def Ensemble():
    from posidonius_amd.ensemble import Ensemble
    return Ensemble


def directions(spec, rng):
    """Eight tangent directions: one per state block, one per parameter, and
    the host's dissipation factor."""
    nb = spec.n_bodies
    B = spec.body
    host = np.zeros(nb, dtype=bool)
    host[spec.body_offset[:-1]] = True
    sig = B["tides_dissipation_factor"] * B["tides_dissipation_factor_scale"]
    d = {k: np.zeros((8, nb, 3)) for k in ("x", "v", "L")}
    d.update({k: np.zeros((8, nb)) for k in ("sigma", "k2", "k2f")})
    d["x"][0] = rng.standard_normal((nb, 3)) * 1e-3 * np.abs(B["position"]).max(axis=1, keepdims=True)
    d["v"][1] = rng.standard_normal((nb, 3)) * 1e-3 * np.abs(B["velocity"]).max(axis=1, keepdims=True)
    d["L"][2] = rng.standard_normal((nb, 3)) * 1e-3 * np.abs(B["angular_momentum"]).max(axis=1, keepdims=True)
    d["sigma"][3] = np.where(host, 0., sig * rng.random(nb))
    d["k2"][4] = np.where(host, 0., B["tides_love_number"] * rng.random(nb))
    d["k2f"][5] = np.where(host, 0., B["flattening_love_number"] * rng.random(nb))
    d["sigma"][6] = np.where(host, sig, 0.)
    d["x"][7] = rng.standard_normal((nb, 3)) * 1e-3 * np.abs(B["position"]).max(axis=1, keepdims=True)
    d["k2"][7] = B["tides_love_number"] * rng.random(nb)
    return d


def per_system(spec, a):
    """Sum of a per-body array over the bodies of each system."""
    a = a.reshape(spec.n_bodies, -1).sum(axis=1)
    return np.add.reduceat(a, spec.body_offset[:-1])


def check_against_tangent(ens, spec, w_rng, d, tan, label, tol=1e-6):
    """Every seed block against every tangent direction: the adjoint's
    g . d must equal w . (J d) to `tol` relative, system by system."""
    fin = ens.get_state()
    for blk in BLOCKS:
        w = w_rng.standard_normal((spec.n_bodies, 3)) / np.abs(fin[blk]).max(axis=1, keepdims=True)
        g = ens.adjoint(**{blk: w})
        tkey = {"position": "x", "velocity": "v", "angular_momentum": "L"}[blk]
        for k in range(8):
            lhs = per_system(spec, w * tan[tkey][k])
            rhs = (per_system(spec, g["position"] * d["x"][k]) + per_system(spec, g["velocity"] * d["v"][k])
                   + per_system(spec, g["angular_momentum"] * d["L"][k])
                   + per_system(spec, g["dissipation_factor"] * d["sigma"][k])
                   + per_system(spec, g["love_number"] * d["k2"][k])
                   + per_system(spec, g["fluid_love_number"] * d["k2f"][k]))
            scale = np.maximum(np.abs(lhs), np.abs(rhs))
            rel = np.abs(lhs - rhs) / np.where(scale > 0, scale, 1.)
            print(f"{label}: seed {blk:17s} dir {k}: max rel {rel.max():.2e}", flush=True)
            assert rel.max() < tol, (blk, k, rel.max(), lhs[:3], rhs[:3])


@pytest.mark.parametrize("steps,segment", [(200, 64), (37, 1000), (10_000, 1000), (100_000, 1000)])
def test_adjoint_equals_tangent_linear(Ensemble, oracle, steps, segment):
    """16 systems of the adjoint workload, up to its full horizon (10^5
    steps with a checkpoint every 1,000): every seed block against every
    tangent direction to 1e-6 relative."""
    spec = gen.config4(16)
    print(f"adjoint vs tangent-linear: {spec.n_systems} systems x {steps} steps, segments of {segment}", flush=True)
    rng = np.random.default_rng(17)
    d = directions(spec, rng)
    ref, tan = oracle.tangent(spec, steps, d, threads=max(1, min(16, os.cpu_count() or 1)))
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=segment)
        fin = ens.get_state()
        for k in ("position", "velocity", "angular_momentum"):
            np.testing.assert_array_equal(fin[k], ref[k], err_msg=f"recorded forward run differs in {k}")
        worst = 0.0
        for bi, blk in enumerate(BLOCKS):
            w = rng.standard_normal((spec.n_bodies, 3)) / np.abs(fin[blk]).max(axis=1, keepdims=True)
            g = ens.adjoint(**{blk: w})
            after = ens.get_state()
            np.testing.assert_array_equal(after["position"], fin["position"])
            tkey = {"position": "x", "velocity": "v", "angular_momentum": "L"}[blk]
            for k in range(8):
                lhs = per_system(spec, w * tan[tkey][k])
                rhs = (per_system(spec, g["position"] * d["x"][k]) + per_system(spec, g["velocity"] * d["v"][k])
                       + per_system(spec, g["angular_momentum"] * d["L"][k])
                       + per_system(spec, g["dissipation_factor"] * d["sigma"][k])
                       + per_system(spec, g["love_number"] * d["k2"][k])
                       + per_system(spec, g["fluid_love_number"] * d["k2f"][k]))
                scale = np.maximum(np.abs(lhs), np.abs(rhs))
                rel = np.abs(lhs - rhs) / np.where(scale > 0, scale, 1.)
                print(f"seed {blk:17s} dir {k}: max rel {rel.max():.2e}  (|w.Jd| ~ {np.median(np.abs(lhs)):.2e})",
                      flush=True)
                worst = max(worst, rel.max())
                assert rel.max() < 1e-6, (blk, k, rel.max(), lhs[:3], rhs[:3])


def wind_workload(n_systems):
    """The adjoint workload with a stellar wind (wind.rs:61-91) on every star.
    The k factor is 1e6 times the reference's solar value (4e-18) so that the
    wind changes the stars' spins measurably over the run and its derivative
    is resolved at the test's tolerance; the even systems' stars spin below
    the saturation rate (1.7593 per day) and the odd ones above it, so both
    branches of the torque are exercised."""
    from posidonius_amd import abi
    spec = gen.config4(n_systems)
    B = spec.body
    nb = spec.n_bodies
    host = np.zeros(nb, dtype=bool)
    host[spec.body_offset[:-1]] = True
    spec.system["effects"] = spec.system["effects"] | abi.EFFECT_WIND
    B["roles"] = np.where(host, B["roles"] | abi.ROLE_WIND, B["roles"]).astype(B["roles"].dtype)
    B["wind_k_factor"] = np.where(host, 4e-12, 0.)
    B["wind_rotation_saturation"] = np.full(nb, 1.7592918860102842)
    I = B["mass"] * B["radius_of_gyration_2"] * B["radius"] ** 2
    for s in range(spec.n_systems):
        h = spec.body_offset[s]
        w = B["spin"][h]
        target = 0.8 if s % 2 == 0 else 4.0
        B["spin"][h] = w * (target / np.linalg.norm(w))
        B["angular_momentum"][h] = I[h] * B["spin"][h]
    return spec


def test_adjoint_with_wind_equals_tangent_linear(Ensemble, oracle):
    """A stellar wind on every star: the adjoint's wind torque term against
    the tangent-linear run, every seed block and direction to 1e-6 relative.
    Without the wind the oracle's tangents of this run differ by up to 12%
    (spin directions) and 5% (position and velocity), so an adjoint that
    dropped or mis-signed the wind term would miss by orders of magnitude."""
    spec = wind_workload(16)
    steps, segment = 300, 100
    rng = np.random.default_rng(23)
    d = directions(spec, rng)
    ref, tan = oracle.tangent(spec, steps, d, threads=max(1, min(16, os.cpu_count() or 1)))
    # the wind changes the stars' spins measurably over the run
    B = spec.body
    h = spec.body_offset[:-1]
    change = np.linalg.norm(ref["angular_momentum"][h] - B["angular_momentum"][h], axis=1) / np.linalg.norm(
        B["angular_momentum"][h], axis=1)
    assert change.min() > 1e-4, change.min()
    print(f"adjoint with wind: stars' spin angular momentum changes by {change.min():.2e}..{change.max():.2e}",
          flush=True)
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=segment)
        fin = ens.get_state()
        for k in BLOCKS:
            np.testing.assert_array_equal(fin[k], ref[k], err_msg=f"recorded forward run differs in {k}")
        check_against_tangent(ens, spec, rng, d, tan, "wind")


def evolving_workload(n_systems, steps):
    """The adjoint workload with evolving stars.  The stars follow synthetic
    evolution tables over the run's own span (the reference's stellar tables
    barely move over a few hundred short steps): the even systems' stars are
    Baraffe1998 (radius only), the odd ones Baraffe2015 (radius and radius
    of gyration), the radius growing by 5-11% and the radius of gyration
    moving by up to 10% within the run.  Without the evolution the oracle's
    tangents of this run differ by up to 30%, so an adjoint that kept the
    initial radii and moments of inertia would miss by orders of magnitude."""
    from posidonius_amd import abi
    from posidonius_amd.case import TABLE_FIELDS
    spec = gen.config4(n_systems)
    B = spec.body
    nb = spec.n_bodies
    spec.system["effects"] = spec.system["effects"] | abi.EFFECT_EVOLUTION
    span = 1.1 * steps * float(np.max(spec.system["time_step"]))
    evo = np.zeros(nb, dtype=np.int32)
    offset = np.zeros(nb + 1, dtype=np.int32)
    tabs = {k: [] for k in TABLE_FIELDS}
    for b in range(nb):
        rows = 0
        s = np.searchsorted(spec.body_offset, b, side="right") - 1
        if b == spec.body_offset[s]:
            evo[b] = abi.EVO_BARAFFE1998 if s % 2 == 0 else abi.EVO_BARAFFE2015
            t = np.array([0., span / 3, 2 * span / 3, span])
            tabs["evolver_time"].append(t)
            tabs["evolver_radius"].append(B["radius"][b] * np.array([1., 1.04, 1.07, 1.12]))
            tabs["evolver_radius_of_gyration_2"].append(B["radius_of_gyration_2"][b] * np.array([1., 1.05, 0.97, 1.1]))
            tabs["evolver_love_number"].append(np.full(4, B["tides_love_number"][b]))
            tabs["evolver_inverse_tidal_q_factor"].append(np.zeros(4))
            rows = 4
        offset[b + 1] = offset[b] + rows
    B["evolution_type"] = evo
    B["evolver_left_index"] = np.zeros(nb, dtype=np.int32)
    spec.evolver_offset = offset
    spec.tables = {k: np.concatenate(v) for k, v in tabs.items()}
    return spec


def test_adjoint_with_evolution_equals_tangent_linear(Ensemble, oracle):
    """Evolving stars: the reverse sweep takes each step's radius and moment
    of inertia from the tape; against the tangent-linear run, every seed
    block and direction to 1e-6 relative."""
    steps, segment = 300, 100
    spec = evolving_workload(16, steps)
    rng = np.random.default_rng(29)
    d = directions(spec, rng)
    ref, tan = oracle.tangent(spec, steps, d, threads=max(1, min(16, os.cpu_count() or 1)))
    h = spec.body_offset[:-1]
    grown = ref["radius"][h] / spec.body["radius"][h]
    assert grown.min() > 1.03, grown.min()
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=segment)
        fin = ens.get_state()
        for k in BLOCKS + ("radius", "moment_of_inertia"):
            np.testing.assert_array_equal(fin[k], ref[k], err_msg=f"recorded forward run differs in {k}")
        check_against_tangent(ens, spec, rng, d, tan, "evolution")


def coordinate_workloads():
    """Systems in WHDS coordinates (whfast.rs:973-1023, 1090-1155, 495-556:
    the star + 3 planet workload, systems of 2-12 bodies, and the three
    coordinate types mixed in wavefronts) and in Jacobi coordinates
    (whfast.rs:889-963, 1034-1080): the
    star + 3 planet workload switched to Jacobi coordinates, the binary host
    with 1-3 circumbinary planets (3-5 bodies: the host / first body pair
    left out of the gravity, the indirect term on the planets), the same
    with each primary moved to a later index (the chain then runs host first
    and the others by index, and the pair left out is the host and the
    lowest other index), and the two mixed with democratic-heliocentric
    systems so that lanes of both kinds of reverse step share wavefronts."""
    from posidonius_amd import abi
    from posidonius_amd.case import EnsembleSpec
    four = gen.config4(16)
    four.system["coordinates"][:] = abi.COORD_JACOBI
    mixed = EnsembleSpec.concat([gen.config4(8, seed=31), gen.config3(8, seed=33), gen.config2(8, seed=35)])
    whds4 = gen.config4(16, seed=39)
    whds4.system["coordinates"][:] = abi.COORD_WHDS
    whds_sizes = gen.config2(24, seed=41, n_min=2, n_max=12)
    whds_sizes.system["coordinates"][:] = abi.COORD_WHDS
    all3 = EnsembleSpec.concat([gen.config4(8, seed=43), gen.config3(8, seed=45), whds4])
    return {"star_3planets": four, "binary_circumbinary": gen.config3(16), "mixed": mixed,
            "host_moved": move_host(gen.config3(16, seed=37), np.random.default_rng(37), first=1)[0],
            "whds_star_3planets": whds4, "whds_2_12_bodies": whds_sizes, "all_three_coordinates": all3}


@pytest.mark.parametrize("name", ["star_3planets", "binary_circumbinary", "mixed", "host_moved", "whds_star_3planets",
                                  "whds_2_12_bodies", "all_three_coordinates"])
@pytest.mark.parametrize("steps,segment", [(300, 100), (3_000, 1000)])
def test_adjoint_in_other_coordinates_equals_tangent_linear(Ensemble, oracle, name, steps, segment):
    """Jacobi coordinates (the reverse step's transforms, Kepler drifts of the
    chain's masses, skipped pair and indirect term) and WHDS coordinates (the
    velocity factors, Kepler drifts about the pair's mass, the weighted jump
    and the scaled interaction kick) against the tangent-linear run, every
    seed block and direction to 1e-6 relative."""
    from posidonius_amd import abi
    spec = coordinate_workloads()[name]
    n_jacobi = int(np.sum(spec.system["coordinates"] == abi.COORD_JACOBI))
    n_whds = int(np.sum(spec.system["coordinates"] == abi.COORD_WHDS))
    print(f"adjoint in other coordinates: {name}, {spec.n_systems} systems ({n_jacobi} Jacobi, {n_whds} WHDS) x {steps} "
          f"steps, segments of {segment}", flush=True)
    rng = np.random.default_rng(41)
    d = directions(spec, rng)
    ref, tan = oracle.tangent(spec, steps, d, threads=max(1, min(16, os.cpu_count() or 1)))
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.adjoint_record(steps, segment_steps=segment)
        fin = ens.get_state()
        for k in BLOCKS:
            np.testing.assert_array_equal(fin[k], ref[k], err_msg=f"recorded forward run differs in {k}")
        check_against_tangent(ens, spec, rng, d, tan, f"coordinates {name}")


def test_adjoint_is_the_same_for_every_launch_mapping(Ensemble):
    """Systems of 2-9 bodies: the recorded run and the adjoint give the same
    bits whether each size has its own wavefronts (exactly N or the next power
    of two lanes per system) or sizes share wavefronts."""
    spec = gen.config2(96, seed=11, n_min=2, n_max=9)
    rng = np.random.default_rng(5)
    w = {k: rng.standard_normal((spec.n_bodies, 3)) for k in BLOCKS}
    out = {}
    for g in (-1, 0, -2):
        print(f"adjoint with group size {g}", flush=True)
        with Ensemble(spec, mapping="packed", group_size=g, steps_per_launch=16) as ens:
            ens.adjoint_record(120, segment_steps=50)
            fin = ens.get_state()
            out[g] = ({k: fin[k] for k in BLOCKS}, ens.adjoint(**w))
    for g in (0, -2):
        for part in range(2):
            for k, a in out[-1][part].items():
                assert np.all(np.isfinite(a)), (k, g)
                np.testing.assert_array_equal(out[g][part][k], a, err_msg=f"group size {g}: {k}")
