"""Historic snapshots written on the device (output.rs:90-134 records, taken
where WHFast::iterate takes them, whfast.rs:240-256)."""
import numpy as np
import pytest

from posidonius_amd import abi, generators as gen

pytestmark = pytest.mark.gpu


def snapshot_steps(t0, dt, period, n_steps, last=-1.0):
    """Steps (counted from the start) at whose beginning the reference takes
    a historic snapshot, replaying its time accumulation."""
    out, t = [], t0
    for k in range(n_steps):
        if last < 0.0 or last + period <= t:
            out.append(k)
            last = 0.0 if last < 0.0 else last + period
        t = t + dt
    return out


def test_history_records_match_the_oracle_state(oracle):
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config1(1)    # non-evolving bodies: the snapshot's spin is L / I with the input I
    spec.system["historic_snapshot_period"][:] = 7.3 * spec.system["time_step"]
    n = 120
    with Ensemble(spec, mapping="packed", steps_per_launch=16) as ens:
        ens.enable_history(64)
        ens.step(n)
        rec = ens.drain_history()
        st = ens.get_state(fields=("status",))
    assert not (st["status"][0] & abi.STATUS_HISTORY_OVERFLOW)
    steps = snapshot_steps(0.0, spec.system["time_step"][0], spec.system["historic_snapshot_period"][0], n)
    nb = spec.n_bodies
    assert len(rec) == len(steps) * nb
    I = spec.body["moment_of_inertia"]
    for r, k in enumerate(steps[:6]):
        ref = oracle.run(spec, k) if k > 0 else None
        got = rec[r * nb:(r + 1) * nb]
        x = spec.body["position"] if ref is None else ref["position"]
        v = spec.body["velocity"] if ref is None else ref["velocity"]
        L = spec.body["angular_momentum"] if ref is None else ref["angular_momentum"]
        np.testing.assert_array_equal(got["position"], x)
        np.testing.assert_array_equal(got["velocity"], v)
        np.testing.assert_allclose(got["spin"], L / I[:, None], rtol=1e-15)
        assert np.all(got["body"] == np.arange(nb))
        np.testing.assert_allclose(got["time"], k * spec.system["time_step"][0], rtol=1e-12)
    planets = rec["body"] > 0
    assert np.all(rec["denergy_dt"][planets] != 0) and np.all(rec["denergy_dt"][~planets] == 0)


def test_history_overflow_is_flagged_not_fatal():
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config1(8)
    spec.system["historic_snapshot_period"][:] = spec.system["time_step"]
    with Ensemble(spec, steps_per_launch=8) as ens:
        ens.enable_history(4)
        ens.step(20)
        st = ens.get_state(fields=("status", "current_iteration"))
        rec = ens.drain_history()
    assert np.all(st["status"] & abi.STATUS_HISTORY_OVERFLOW)
    assert np.all(st["current_iteration"] == 20)
    assert len(rec) == 4 * spec.n_bodies


def test_history_with_graph_replay_equals_direct_launches():
    """Enabling the history changes the kernels' arguments, so a handle that
    replays hipGraphs recaptures them: records and state equal those of
    direct launches, bit for bit."""
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config2(64)
    spec.system["historic_snapshot_period"][:] = 5.5 * spec.system["time_step"]
    out = {}
    for graph in (False, True):
        with Ensemble(spec, mapping="packed", group_size=-1, steps_per_launch=8, use_graph=graph) as ens:
            ens.step(8)   # before the history is enabled (with a graph: captured at creation)
            ens.enable_history(32)
            ens.step(64)  # two replays of four 8-step launches
            out[graph] = (ens.drain_history(), ens.get_state(fields=("position", "velocity", "angular_momentum")))
    rec0, st0 = out[False]
    rec1, st1 = out[True]
    assert len(rec0) == len(rec1) > 0
    for name in rec0.dtype.names:
        np.testing.assert_array_equal(rec1[name], rec0[name])
    for k in st0:
        np.testing.assert_array_equal(st1[k], st0[k])


def _records_equal(a, b):
    assert len(a) == len(b) > 0
    for name in a.dtype.names:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


def test_tuning_appends_no_history_records():
    """pos_tune probes the launch mappings on a copy of the state; the probe
    steps are not part of the run, so the history holds only the caller's
    steps and equals that of an untuned handle."""
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config2(256, seed=5)
    spec.system["historic_snapshot_period"][:] = 3.5 * spec.system["time_step"]
    out = {}
    for tuned in (False, True):
        with Ensemble(spec, mapping="packed", group_size=-1, steps_per_launch=16, use_graph=True) as ens:
            ens.enable_history(64)
            ens.step(10)
            if tuned:
                ens.tune(probe_steps=16)
            ens.step(70)
            out[tuned] = (ens.drain_history(), ens.get_state(fields=("position", "status")))
    _records_equal(out[True][0], out[False][0])
    np.testing.assert_array_equal(out[True][1]["position"], out[False][1]["position"])
    assert not np.any(out[True][1]["status"] & abi.STATUS_HISTORY_OVERFLOW)


def test_adjoint_appends_no_history_records():
    """pos_adjoint recomputes every recorded segment from its checkpoint;
    those steps were recorded once already and add no history records."""
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config4(32)
    spec.system["historic_snapshot_period"][:] = 9.5 * spec.system["time_step"]
    rng = np.random.default_rng(3)
    w = rng.standard_normal((spec.n_bodies, 3))
    with Ensemble(spec, mapping="packed", steps_per_launch=16) as ens:
        ens.enable_history(64)
        ens.step(200)
        plain = ens.drain_history()
    with Ensemble(spec, mapping="packed", steps_per_launch=16) as ens:
        ens.enable_history(64)
        ens.adjoint_record(200, segment_steps=50)
        ens.adjoint(position=w)
        adj = ens.drain_history()
        after = ens.get_state(fields=("status",))
    _records_equal(adj, plain)
    assert not np.any(after["status"] & abi.STATUS_HISTORY_OVERFLOW)
