"""Parity of the HIP path with the CPU oracle and the reference's fixtures.

Every comparison here is bitwise: the device code and the oracle perform the
same IEEE binary64 operations in the same order (both built without
multiply-add contraction), so a single differing bit is a failure.
"""
import os
import sys

import numpy as np
import pytest

from posidonius_amd import abi, generators as gen
from posidonius_amd.case import EnsembleSpec
from tests.oracle_runner import golden_names, load_golden
from tests.spec_tools import move_host

pytestmark = pytest.mark.gpu

SUPPORTED = golden_names()
THREADS = max(1, min(16, os.cpu_count() or 1))   # host threads for the oracle
STATE_KEYS = ("position", "velocity", "acceleration", "angular_momentum", "spin", "velocity_error",
              "angular_momentum_error", "radius", "moment_of_inertia", "current_time", "current_iteration",
              "status", "evolver_left_index", "timestep_warning", "pair_dissipation_factor")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def Ensemble():
    from posidonius_amd.ensemble import Ensemble
    return Ensemble


def assert_same_state(gpu, ref, keys=STATE_KEYS, where=""):
    for k in keys:
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        if not np.array_equal(a, b, equal_nan=a.dtype.kind == "f"):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            diff = np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))
            raise AssertionError(f"{where} {k}: {len(bad)} entries differ, first at {i}: gpu={a[i]!r} "
                                 f"oracle={b[i]!r}, max abs diff {diff:.3e}")


def test_fixtures_one_ensemble_bitwise(Ensemble):
    """All 26 reference fixtures in one packed ensemble, run to
    completion (199 steps): positions, velocities and accelerations equal the
    reference's particle_<i>.json bit for bit."""
    specs, expected = [], []
    for name in SUPPORTED:
        s, e = load_golden(name)
        specs.append(s)
        expected.append(e)
    spec = EnsembleSpec.concat(specs)
    log(f"fixtures ensemble: {spec.n_systems} systems, {spec.n_bodies} bodies")
    with Ensemble(spec, mapping="packed") as ens:
        ens.step(250)
        out = ens.get_state()
    assert np.all(out["current_iteration"] == 199)
    assert np.all(out["status"] & abi.STATUS_COMPLETED)
    for s, name in enumerate(SUPPORTED):
        b0, b1 = spec.body_offset[s], spec.body_offset[s + 1]
        for k in ("position", "velocity", "acceleration"):
            np.testing.assert_array_equal(out[k][b0:b1], expected[s][k], err_msg=f"{name} {k}")


@pytest.mark.parametrize("mapping", ["plain", "packed"])
def test_fixtures_against_oracle_full_state(Ensemble, oracle, mapping):
    names = SUPPORTED[::4]
    spec = EnsembleSpec.concat([load_golden(n)[0] for n in names])
    ref = oracle.run(spec, 120)
    log(f"fixtures vs oracle mapping={mapping}: {spec.n_systems} systems x 120 steps")
    with Ensemble(spec, mapping=mapping) as ens:
        ens.step(120)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"[{mapping}]")


@pytest.mark.parametrize("cfg,n,steps", [(0, None, 1000), (1, 256, 1000), (2, 1024, 1000), (3, 512, 1000), (4, 64, 1000)])
def test_configs_bitwise_against_oracle(Ensemble, oracle, cfg, n, steps):
    """Every benchmark configuration at the 10^3-step horizon of the parity
    target (1e-10 relative); the device meets it bit for bit."""
    spec = gen.CONFIGS[cfg]() if n is None else gen.CONFIGS[cfg](n)
    log(f"config {cfg}: {spec.n_systems} systems x {steps} steps")
    ref = oracle.run(spec, steps, threads=THREADS)
    with Ensemble(spec, mapping="packed", steps_per_launch=64) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"config {cfg}")


@pytest.mark.parametrize("group_size", [-2, -1, 0])
def test_jacobi_fixed_effect_set_all_sizes(Ensemble, oracle, group_size):
    """The fixed-effect-set kernel in Jacobi coordinates on systems of 2 to
    16 bodies (two bodies: the one pair is the one the splitting leaves out of
    the gravity and no body gets the indirect term), mixed in wavefronts or in
    size classes: bitwise equal to the oracle."""
    spec = gen.config2(256, seed=47, n_min=2, n_max=abi.MAX_BODIES)
    spec.system["coordinates"][:] = abi.COORD_JACOBI
    steps = 200
    log(f"Jacobi fixed effect set: {spec.n_systems} systems of 2-{abi.MAX_BODIES} bodies x {steps} steps, G={group_size}")
    ref = oracle.run(spec, steps, threads=THREADS)
    with Ensemble(spec, mapping="packed", group_size=group_size, steps_per_launch=64) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"Jacobi fixed effect set G={group_size}")


@pytest.mark.parametrize("group_size,K,graph", [(0, 1, False), (0, 7, False), (8, 64, True), (16, 32, False), (3, 16, True), (-1, 64, True),
                                               (-2, 64, True), (-2, 16, False)])
def test_launch_mappings_agree(Ensemble, oracle, group_size, K, graph):
    """Group size, steps per launch and graph replay do not change a bit."""
    spec = gen.config2(96) if group_size in (16, -1) else gen.config1(96)
    if group_size == 3:
        spec = gen.config1(96)
    if group_size == -2:   # sizes 2-16 mixed in wavefronts
        spec = gen.config2(256, seed=43, n_min=2, n_max=abi.MAX_BODIES)
    steps = 100
    ref = oracle.run(spec, steps, threads=THREADS)
    log(f"mapping G={group_size} K={K} graph={graph}")
    with Ensemble(spec, mapping="packed", group_size=group_size, steps_per_launch=K, use_graph=graph) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"G={group_size} K={K} graph={graph}")


@pytest.mark.parametrize("mapping,group_size,graph", [("packed", 0, False), ("plain", 0, False), ("packed", -2, True)])
def test_operators_against_oracle(Ensemble, oracle, mapping, group_size, graph):
    """pos_gravity_acceleration / pos_additional_effects: the two operator-level
    entry points the reference benchmarks (benches/simulation.rs:127, 138),
    over each launch table; steps taken after them are unaffected."""
    spec = EnsembleSpec.concat([gen.config2(64), gen.config3(16), load_golden("test_integrator-whfast_jacobi")[0]])
    log(f"operators over {mapping} G={group_size} graph={graph}")
    with Ensemble(spec, mapping=mapping, group_size=group_size, steps_per_launch=16, use_graph=graph) as ens:
        g_acc, g_st = ens.gravity_acceleration()
        e_acc, e_dl = ens.additional_effects()
        ens.step(64)
        after = ens.get_state()
    assert_same_state(after, oracle.run(spec, 64, threads=8), where=f"steps after the operators ({mapping})")
    r_acc, r_st = oracle.gravity(spec)
    np.testing.assert_array_equal(g_acc, r_acc)
    np.testing.assert_array_equal(g_st, r_st)
    r_eacc, r_edl = oracle.additional_effects(spec)
    np.testing.assert_array_equal(e_acc, r_eacc)
    np.testing.assert_array_equal(e_dl, r_edl)


def test_operator_per_effect_set_on_the_benchmark_universe(Ensemble, oracle):
    """pos_additional_effects on the universe of the reference's operator
    benchmark (benches/simulation.rs:14-46, 138-221: a star with a disk and
    GalletBolmont2017 evolution, four planets), once per effect set of that
    benchmark, 8 copies each, against the oracle bit for bit; bench.py times
    the same sets."""
    import bench
    base = bench.benchmark_universe()
    spec = EnsembleSpec.concat([base] * 8)
    for name, mask in bench.EFFECT_SETS:
        s = spec.copy()
        s.system["effects"] = s.system["effects"].copy()
        s.system["effects"][:] = mask
        log(f"effect set {name}")
        with Ensemble(s, mapping="packed", group_size=-1) as ens:
            acc, dl = ens.additional_effects()
        r_acc, r_dl = oracle.additional_effects(s)
        np.testing.assert_array_equal(acc, r_acc, err_msg=name)
        np.testing.assert_array_equal(dl, r_dl, err_msg=name)
        if mask & (abi.EFFECT_TIDES | abi.EFFECT_DISK | abi.EFFECT_ROTATIONAL_FLATTENING | abi.EFFECT_GENERAL_RELATIVITY):
            assert np.abs(acc).max() > 0, name


def test_resume_equals_continuous_run(Ensemble, oracle):
    """get_state -> new ensemble -> continue gives the same bits as one run
    (checkpoint/resume, src/integrator/output.rs:28-54, 178-203)."""
    spec = gen.config1(64)
    with Ensemble(spec) as ens:
        ens.step(300)
        mid = ens.get_state()
        ens.step(200)
        full = ens.get_state()
    resumed = spec.copy()
    for k in ("position", "velocity", "angular_momentum", "spin", "velocity_error", "angular_momentum_error"):
        resumed.body[k] = mid[k].copy()
    resumed.system["current_time"] = mid["current_time"].copy()
    resumed.system["last_historic_snapshot_time"] = mid["last_historic_snapshot_time"].copy()
    resumed.system["current_iteration"] = mid["current_iteration"].copy()
    with Ensemble(resumed) as ens2:
        ens2.set_state({"acceleration": mid["acceleration"]})
        ens2.step(200)
        out = ens2.get_state()
    assert_same_state(out, full, keys=("position", "velocity", "angular_momentum", "spin", "velocity_error",
                                       "angular_momentum_error", "current_time", "current_iteration"))


def test_collision_freezes_system(Ensemble, oracle):
    """A collision (universe.rs:230-233) raises POS_STATUS_COLLISION and
    freezes that system only; the oracle agrees bit for bit."""
    spec = gen.config1(8)
    b = spec.body_offset[3] + 1
    spec.system["effects"][3] = 0          # point masses: the radius enters only the collision test
    spec.body["radius"][b] = 0.2          # inflate one planet past the host
    ref = oracle.run(spec, 50)
    with Ensemble(spec) as ens:
        ens.step(50)
        out = ens.get_state()
    assert out["status"][3] & abi.STATUS_COLLISION
    assert out["current_iteration"][3] == 1
    assert np.all(out["current_iteration"][np.arange(8) != 3] == 50)
    assert_same_state(out, ref, where="collision")


def test_benchmark_ensemble_at_horizon(Ensemble, oracle):
    """The benchmark ensemble itself (65,536 systems of 2-12 bodies) for 10^3
    steps: every system equals the oracle bit for bit under the benchmark's
    launch shape (exactly N lanes per system, 64 steps per launch, graph
    replay), and a second mapping (16 lanes per system, 16 steps per launch,
    direct launches) gives the same bits."""
    spec = gen.config2(65536)
    steps = 1000
    log(f"benchmark ensemble: {spec.n_systems} systems x {steps} steps")
    with Ensemble(spec, group_size=-1, steps_per_launch=64, use_graph=True) as ens:
        ens.step(steps)
        out = ens.get_state()
    with Ensemble(spec, group_size=16, steps_per_launch=16) as ens:
        ens.step(steps)
        out2 = ens.get_state()
    assert_same_state(out2, out, where="determinism")
    log(f"oracle on {THREADS} threads")
    ref = oracle.run(spec, steps, threads=THREADS)
    assert_same_state(out, ref, where="benchmark ensemble")
    log(f"{int(np.count_nonzero(out['current_iteration'] < steps))} systems stopped early (fatal status)")


@pytest.mark.parametrize("cfg,n", [(1, 256), (2, 512)])
def test_configs_at_their_horizon(Ensemble, oracle, cfg, n):
    """The uniform (3 bodies) and the uneven (2-12 bodies) ensembles over
    their full 10^4 steps, on a subset of their systems, under the
    benchmark's launch shape: every system equals the oracle bit for bit."""
    spec = gen.CONFIGS[cfg](n)
    steps = 10_000
    log(f"config {cfg}: {spec.n_systems} systems x {steps} steps")
    with Ensemble(spec, mapping="packed", group_size=-2, steps_per_launch=64, use_graph=True) as ens:
        ens.step(steps)
        out = ens.get_state()
    ref = oracle.run(spec, steps, threads=THREADS)
    assert_same_state(out, ref, where=f"config {cfg} at 1e4 steps")


def test_binary_hosts_at_their_horizon(Ensemble, oracle):
    """The binary-host workload (2 stars + 1-3 circumbinary planets, Jacobi
    coordinates, the Jacobi fixed-effect-set kernel) over its full 10^5
    steps, on 64 of its systems: every system equals the oracle bit for bit,
    status and Kepler-solver flags included."""
    spec = gen.config3(64)
    steps = 100_000
    log(f"binary hosts: {spec.n_systems} systems x {steps} steps")
    with Ensemble(spec, mapping="packed", steps_per_launch=64, use_graph=True) as ens:
        ens.step(steps)
        out = ens.get_state()
    log(f"oracle on {THREADS} threads")
    ref = oracle.run(spec, steps, threads=THREADS)
    assert_same_state(out, ref, where="binary hosts at 1e5 steps")
    log(f"status bits set: {sorted(set(int(x) for x in out['status']))}")


def gr_variant_ensemble():
    """The uneven-size ensemble and the Jacobi binary-host ensemble, each under
    every general-relativity implementation (Kidder, Anderson, Newhall,
    disabled), interleaved so that device order differs from input order."""
    parts = []
    for impl in (abi.GR_KIDDER1995, abi.GR_ANDERSON1975, abi.GR_NEWHALL1983, abi.GR_DISABLED):
        for s in (gen.config2(24, seed=11 + impl), gen.config3(8, seed=21 + impl)):
            s.system["gr_implementation"][:] = impl
            parts.append(s)
    spec = EnsembleSpec.concat(parts)
    return spec.subset(np.random.default_rng(5).permutation(spec.n_systems))


@pytest.mark.parametrize("mapping,group_size,graph", [("packed", -1, True), ("packed", 0, False), ("plain", 0, False),
                                                      ("packed", -2, False)])
def test_gr_variants_against_oracle(Ensemble, oracle, mapping, group_size, graph):
    """Anderson 1975 and Newhall 1983 (general_relativity.rs:461-895) run in
    launches of their own; all four implementations in one ensemble agree
    with the oracle bit for bit."""
    spec = gr_variant_ensemble()
    steps = 80
    ref = oracle.run(spec, steps, threads=THREADS)
    log(f"GR variants mapping={mapping} G={group_size} graph={graph}: {spec.n_systems} systems x {steps} steps")
    with Ensemble(spec, mapping=mapping, group_size=group_size, steps_per_launch=16, use_graph=graph) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"GR variants {mapping} G={group_size}")


def test_gr_variants_operators_against_oracle(Ensemble, oracle):
    spec = gr_variant_ensemble()
    with Ensemble(spec) as ens:
        e_acc, e_dl = ens.additional_effects()
    r_eacc, r_edl = oracle.additional_effects(spec)
    np.testing.assert_array_equal(e_acc, r_eacc)
    np.testing.assert_array_equal(e_dl, r_edl)


def test_adjoint_rejects_other_gr_variants(Ensemble):
    spec = gen.config2(4, seed=3)
    spec.system["gr_implementation"][:] = abi.GR_NEWHALL1983
    with Ensemble(spec) as ens:
        with pytest.raises(Exception, match="Kidder1995"):
            ens.adjoint_record(10, 5)


def test_adjoint_rejects_the_disk(Ensemble):
    """The adjoint refuses the disk with a clear error."""
    spec, _ = load_golden("test_disk-enabled_disk")
    with Ensemble(spec) as ens:
        with pytest.raises(Exception, match="does not support the disk"):
            ens.adjoint_record(10, 5)


def excited_dynamical_tide():
    """The BolmontMathis2016 fixture with a hot-Jupiter-mass first planet and
    the inverse tidal Q factor raised 1e4 times: the star's dynamical tide is
    excited (constant_time_lag.rs:57-64) and the pair factor moves the state
    by ~2e-14 AU within 199 steps."""
    spec, _ = load_golden("test_evolution-solar_like_bolmontmathis2016")
    s = spec.copy()
    for k in ("mass", "moment_of_inertia"):
        s.body[k][1] *= 300.0
    s.body["angular_momentum"][1] *= 300.0
    s.tables["evolver_inverse_tidal_q_factor"] = s.tables["evolver_inverse_tidal_q_factor"] * 1e4
    return s


def excited_orbiting_dynamical_tide():
    """The BolmontMathis2016 fixture whose third body (LeconteChabrier2013)
    has dynamical tides with an inverse tidal Q factor of 0.1: its spin is
    excited by its orbit and its own pair factor (constant_time_lag.rs:96-165)
    becomes ~290 times its scaled factor."""
    spec, _ = load_golden("test_evolution-solar_like_bolmontmathis2016")
    s = spec.copy()
    s.body["evolution_type"][2] = abi.EVO_LECONTE_CHABRIER2013_DYN
    b0, b1 = s.evolver_offset[2], s.evolver_offset[3]
    s.tables["evolver_inverse_tidal_q_factor"][b0:b1] = 0.1
    return s


def extended_ensemble():
    """Systems of the extended effects pass (the disk and evolving dissipation)
    among plain ones: the three reference fixtures, the disk fixture with its
    surface density lowered 1e7 times (below the min(1, 2 Sigma a^2 / m)
    saturation, so the surface density and both exponentials enter), the
    excited dynamical tide, and uneven / Jacobi systems; shuffled."""
    parts = [load_golden(n)[0] for n in ("test_disk-enabled_disk", "test_evolution-solar_like_bolmontmathis2016",
                                         "test_evolution-solar_like_galletbolmont2017")]
    thin = parts[0].copy()
    thin.system["disk_properties"][0, 4] *= 1e-7
    parts += [thin, excited_dynamical_tide(), excited_orbiting_dynamical_tide(), gen.config2(16, seed=31),
              gen.config3(8, seed=32)]
    spec = EnsembleSpec.concat(parts)
    return spec.subset(np.random.default_rng(9).permutation(spec.n_systems))


@pytest.mark.parametrize("mapping,group_size,graph", [("packed", -1, True), ("packed", 0, False), ("plain", 0, False),
                                                      ("packed", -2, True)])
def test_extended_effects_against_oracle(Ensemble, oracle, mapping, group_size, graph):
    """Disk forces (disk.rs:155-268), evolving-dissipation lag angles and pair
    factors (evolution.rs:548-567, constant_time_lag.rs:20-94) in their own
    kernel variant: bitwise equal to the oracle."""
    spec = extended_ensemble()
    steps = 150
    ref = oracle.run(spec, steps, threads=THREADS)
    log(f"extended effects mapping={mapping} G={group_size} graph={graph}: {spec.n_systems} systems x {steps} steps")
    with Ensemble(spec, mapping=mapping, group_size=group_size, steps_per_launch=16, use_graph=graph) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"extended {mapping} G={group_size}")


def test_extended_effects_operators_against_oracle(Ensemble, oracle):
    spec = extended_ensemble()
    with Ensemble(spec) as ens:
        e_acc, e_dl = ens.additional_effects()
    r_eacc, r_edl = oracle.additional_effects(spec)
    np.testing.assert_array_equal(e_acc, r_eacc)
    np.testing.assert_array_equal(e_dl, r_edl)


def test_history_carries_lag_angle_and_migration_timescale(Ensemble):
    """output.rs:131-134: the lag angle of the evolving-dissipation star
    (3 |Omega|^2 / SUN_DYN_FREQ_2 * Q^-1(t) / 4, evolution.rs:553-560) and the
    disk migration timescale of the disk's planets."""
    names = ("test_evolution-solar_like_bolmontmathis2016", "test_disk-enabled_disk")
    spec = EnsembleSpec.concat([load_golden(n)[0] for n in names])
    spec.system["historic_snapshot_period"][:] = 10 * spec.system["time_step"]
    with Ensemble(spec, steps_per_launch=8) as ens:
        ens.enable_history(32)
        ens.step(60)
        rec = ens.drain_history()
    bm = rec[rec["system"] == 0]
    star = bm[bm["body"] == 0]
    assert len(star) == 6 and np.all(star["lag_angle"] > 0)
    assert np.all(bm["lag_angle"][bm["body"] > 0] == 0)
    K2 = (6.67428e-11 / 1.49597870700e11 ** 3) * 1.9818e30 * 86400.0 ** 2
    sun_dyn_freq_2 = K2 / (6.957e8 / 1.49597870700e11) ** 3
    b0 = spec.evolver_offset[0]
    n0 = spec.evolver_offset[1] - b0
    t_tab = spec.tables["evolver_time"][b0:b0 + n0]
    q_tab = spec.tables["evolver_inverse_tidal_q_factor"][b0:b0 + n0]
    omega2 = (star["spin"] ** 2).sum(axis=1)
    expected = 3.0 * omega2 / sun_dyn_freq_2 * np.interp(star["time"], t_tab, q_tab) / 4.0
    np.testing.assert_allclose(star["lag_angle"], expected, rtol=1e-12)
    disk = rec[rec["system"] == 1]
    planets = disk[disk["body"] > 0]
    first = planets["time"] == 0.0   # before the first evaluation
    assert np.all(planets["migration_timescale"][first] == 0)
    assert np.all(np.isfinite(planets["migration_timescale"][~first]))
    assert np.all(planets["migration_timescale"][~first] > 0)
    assert np.all(disk["migration_timescale"][disk["body"] == 0] == 0)
    assert np.all(bm["migration_timescale"] == 0)


@pytest.mark.parametrize("name", ["test_integrator-whfast_jacobi", "test_integrator-whfast_democraticheliocentric",
                                  "test_integrator-whfast_whds"])
def test_order_invariance_star_swapped(Ensemble, name):
    """tests/test_order.rs:341-422: the reference's star-swapped cases end on
    the same states at 1e-16; on the device, in one ensemble with the
    unswapped case and under the packed mapping (the host lane is then body 1
    of its group)."""
    from tests.test_oracle_fixtures import _swap_first_two
    spec, expected = load_golden(name)
    swapped = _swap_first_two(spec)
    both = EnsembleSpec.concat([spec, swapped])
    with Ensemble(both, mapping="packed") as ens:
        ens.step(250)
        out = ens.get_state()
    n = spec.n_bodies
    perm = np.array([1, 0] + list(range(2, n)))
    for key in ("position", "velocity", "acceleration"):
        np.testing.assert_array_equal(out[key][:n], expected[key])
        np.testing.assert_allclose(out[key][n:][perm], expected[key], rtol=0, atol=1e-16)


def test_resume_carries_the_pair_dependent_map(Ensemble):
    """An evolving-dissipation run stopped, rebuilt from its state (including
    the orbiting body's pair-dependent map entry and the evolvers'
    left_index) and continued equals the uninterrupted run bit for bit."""
    spec = excited_orbiting_dynamical_tide()
    with Ensemble(spec) as ens:
        ens.step(100)
        mid = ens.get_state()
        ens.step(99)
        full = ens.get_state()
    assert not np.isnan(mid["pair_dissipation_factor"][2])
    resumed = spec.copy()
    for k in ("position", "velocity", "angular_momentum", "spin", "velocity_error", "angular_momentum_error",
              "radius", "radius_of_gyration_2", "moment_of_inertia", "evolver_left_index", "pair_dissipation_factor"):
        resumed.body[k] = mid[k].copy()
    for k in ("current_time", "last_historic_snapshot_time", "current_iteration"):
        resumed.system[k] = mid[k].copy()
    with Ensemble(resumed) as ens2:
        ens2.set_state({"acceleration": mid["acceleration"]})
        ens2.step(99)
        out = ens2.get_state()
    assert_same_state(out, full, keys=("position", "velocity", "angular_momentum", "spin", "velocity_error",
                                       "angular_momentum_error", "current_iteration", "evolver_left_index",
                                       "pair_dissipation_factor"))


def test_ejection_and_close_encounter_freeze_their_systems(Ensemble, oracle):
    """The panics of universe.rs:224-237 other than the collision: a body
    beyond 100 AU from the host (ejection) and a body inside the Roche radius
    (close encounter) set their status bits, freeze only their own systems,
    and match the oracle bit for bit."""
    spec = gen.config1(8)
    spec.system["effects"][[2, 5]] = 0           # point masses
    host2, b2 = spec.body_offset[2], spec.body_offset[2] + 1
    spec.body["position"][b2] = spec.body["position"][host2] + np.array([150.0, 0.0, 0.0])
    host5, b5 = spec.body_offset[5], spec.body_offset[5] + 2
    spec.body["position"][b5] = spec.body["position"][host5] + np.array([0.002, 0.0, 0.0])
    ref = oracle.run(spec, 30)
    with Ensemble(spec) as ens:
        ens.step(30)
        out = ens.get_state()
    assert out["status"][2] & abi.STATUS_EJECTION
    assert out["status"][5] & abi.STATUS_ROCHE
    others = np.setdiff1d(np.arange(8), [2, 5])
    assert np.all(out["current_iteration"][others] == 30)
    assert np.all(out["current_iteration"][[2, 5]] == 1)
    assert_same_state(out, ref, where="ejection / close encounter")


@pytest.mark.parametrize("group_size", [-1, 0, -2])
def test_largest_systems_against_oracle(Ensemble, oracle, group_size):
    """Systems of the maximum size (POS_MAX_BODIES = 16 bodies: four per
    wavefront) under both lane-group mappings."""
    spec = gen.config2(24, seed=41, n_min=abi.MAX_BODIES, n_max=abi.MAX_BODIES)
    ref = oracle.run(spec, 100, threads=8)
    with Ensemble(spec, mapping="packed", group_size=group_size, steps_per_launch=32) as ens:
        ens.step(100)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"16 bodies G={group_size}")


@pytest.mark.parametrize("mapping,group_size", [("packed", -1), ("packed", 0), ("plain", 0), ("packed", -2)])
def test_fixed_effect_set_with_the_host_anywhere(Ensemble, oracle, mapping, group_size):
    """The fixed-effect-set kernel (tides + flattening + Kidder GR in
    heliocentric coordinates, pair terms hoisted out of the implicit-midpoint
    iterations) with the star at any index of its system, so that the host
    lane is not the first lane of its group: bitwise equal to the oracle on
    the same permuted systems."""
    rng = np.random.default_rng(2024)
    spec, perm = move_host(gen.config2(384), rng)
    assert np.any(spec.body["mass"][spec.body_offset[:-1]] < 0.05), "some star must have moved"
    steps = 200
    log(f"host anywhere: {spec.n_systems} systems x {steps} steps, {mapping} G={group_size}")
    ref = oracle.run(spec, steps, threads=THREADS)
    with Ensemble(spec, mapping=mapping, group_size=group_size, steps_per_launch=64) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"host anywhere ({mapping}, G={group_size})")


@pytest.mark.parametrize("n_max", [12, 16])
def test_tune_restores_the_state_and_keeps_the_bits(Ensemble, oracle, n_max):
    """pos_tune's trial launches run on the ensemble's own state and put it
    back; the mapping it chooses continues to the oracle's bits. With sizes up
    to 16 the size classes leave over a tenth of the lanes idle, so the mixed
    packing is among the trials."""
    spec = gen.config2(192, seed=7, n_min=2, n_max=n_max)
    steps = 60
    ref = oracle.run(spec, steps, threads=THREADS)
    log(f"tune: {spec.n_systems} systems of 2-{n_max} bodies")
    with Ensemble(spec, mapping="packed", use_graph=True) as ens:
        ens.step(20)
        before = ens.get_state()
        o = ens.tune(probe_steps=16)
        log(f"tune chose G={o.group_size} K={o.steps_per_launch} graph={o.use_graph}")
        assert o.mapping == abi.MAPPING_PACKED and o.steps_per_launch in (16, 64)
        assert_same_state(ens.get_state(), before, where="after tune")
        ens.step(steps - 20)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"tuned G={o.group_size} K={o.steps_per_launch}")


def test_resume_keeps_the_roche_radii_of_t0(Ensemble, oracle):
    """Roche radii are computed once, at t = 0 (initialize_physical_values,
    whfast.rs:226-233), and a resumed run keeps them (they travel in the
    snapshot, universe.rs:62).  Planet 2 of the fixture has an evolving
    radius, so radii recomputed at the resume would differ from them."""
    spec, _ = load_golden("test_evolution-brown_dwarf_leconte2011")
    spec.system["time_limit"][:] = 1e9
    with Ensemble(spec) as ens:
        r0 = ens.get_state(fields=("roche_radius", "radius"))
        ens.step(150)
        mid = ens.get_state()
        ens.step(150)
        full = ens.get_state()
    np.testing.assert_array_equal(mid["roche_radius"], r0["roche_radius"])
    assert not np.array_equal(mid["radius"], r0["radius"]), "the fixture's radii must evolve"
    resumed = spec.copy()
    for k in ("position", "velocity", "angular_momentum", "spin", "velocity_error", "angular_momentum_error",
              "radius", "radius_of_gyration_2", "moment_of_inertia", "evolver_left_index", "roche_radius"):
        resumed.body[k] = mid[k].copy()
    for k in ("current_time", "last_historic_snapshot_time", "current_iteration"):
        resumed.system[k] = mid[k].copy()
    with Ensemble(resumed) as ens2:
        np.testing.assert_array_equal(ens2.get_state(fields=("roche_radius",))["roche_radius"], r0["roche_radius"])
        ens2.set_state({"acceleration": mid["acceleration"]})
        ens2.step(150)
        out = ens2.get_state()
    assert_same_state(out, full, keys=("position", "velocity", "angular_momentum", "spin", "radius",
                                       "roche_radius", "current_time", "current_iteration"))
    # the oracle keeps the carried radii too
    ref = oracle.run(resumed, 150)
    np.testing.assert_array_equal(ref["roche_radius"], r0["roche_radius"])
    # without them, a resumed system gets radii from its current radii
    resumed.body["roche_radius"] = None
    with Ensemble(resumed) as ens3:
        r3 = ens3.get_state(fields=("roche_radius",))["roche_radius"]
    assert not np.array_equal(r3, r0["roche_radius"])
