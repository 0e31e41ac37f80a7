"""`posidonius-hip start|resume` end to end on the GPU.

A run to completion and a run that stops early (shorter time limit) and is
resumed from its last recovery snapshot with the full time limit must write
byte-identical history files (the resume truncates the history to what the
snapshot accounts for, output.rs:64-87, and continues bit for bit).  The
history records follow the reference's 156-byte layout (output.rs:90-134).
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from tests.oracle_runner import GOLDEN_DIR

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "posidonius_amd", "posidonius-hip")
RECORD = struct.Struct("<ddi12d5d")


def read_history(path, n):
    data = open(path, "rb").read()
    assert len(data) % (RECORD.size * n) == 0
    return [RECORD.unpack_from(data, o) for o in range(0, len(data), RECORD.size)]


def write_case(path, hist_every, rec_every, time_limit=None):
    with open(os.path.join(GOLDEN_DIR, "test_integrator-whfast_democraticheliocentric.json")) as f:
        case = json.load(f)["case"]
    dt = case["time_step"]
    case["historic_snapshot_period"] = hist_every * dt
    case["recovery_snapshot_period"] = rec_every * dt
    if time_limit is not None:
        case["universe"]["time_limit"] = time_limit
    with open(path, "w") as f:
        json.dump(case, f)
    return case


def cli(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r


def test_start_and_resume_write_identical_history(tmp_path):
    full = write_case(tmp_path / "case.json", 10, 50)
    n = full["universe"]["n_particles"]
    cli("start", str(tmp_path / "case.json"), str(tmp_path / "a.json"), str(tmp_path / "a.bin"), "-s")
    rec_a = read_history(tmp_path / "a.bin", n)
    # first snapshot: the initial state, ids and times in order
    first = rec_a[:n]
    for i, r in enumerate(first):
        assert r[0] == 0.0 and r[2] == i
        p = full["universe"]["particles"][i]["inertial_position"]
        assert r[3:6] == (p["x"], p["y"], p["z"])
    times = [r[0] for r in rec_a[::n]]
    assert times == sorted(times) and len(times) >= 19
    # stop early, then resume with the full time limit
    write_case(tmp_path / "short.json", 10, 50, time_limit=full["universe"]["time_limit"] * 0.6)
    cli("start", str(tmp_path / "short.json"), str(tmp_path / "b.json"), str(tmp_path / "b.bin"), "-s")
    assert os.path.getsize(tmp_path / "b.bin") < os.path.getsize(tmp_path / "a.bin")
    snap = json.load(open(tmp_path / "b.json"))
    assert snap["current_iteration"] > 0 and snap["n_historic_snapshots"] > 0
    cli("resume", str(tmp_path / "b.json"), str(tmp_path / "b.bin"), "-s", "--time-limit",
        repr(full["universe"]["time_limit"]))
    assert open(tmp_path / "b.bin", "rb").read() == open(tmp_path / "a.bin", "rb").read()


@pytest.mark.parametrize("name", ["test_disk-enabled_disk", "test_evolution-solar_like_bolmontmathis2016",
                                  "test_general_relativity-newhall1983"])
def test_start_history_equals_the_oracle(tmp_path, oracle, name):
    """A reference case run to its time limit by the command-line host: every
    historic snapshot (taken at the start of steps 0, 20, ..., 180) holds the
    oracle's state after that many steps, bit for bit, and the records carry
    the disk migration timescale and the star's lag angle."""
    from posidonius_amd.case import read_case
    with open(os.path.join(GOLDEN_DIR, name + ".json")) as f:
        case = json.load(f)["case"]
    n = case["universe"]["n_particles"]
    case["historic_snapshot_period"] = 20 * case["time_step"]
    with open(tmp_path / "case.json", "w") as f:
        json.dump(case, f)
    cli("start", str(tmp_path / "case.json"), str(tmp_path / "s.json"), str(tmp_path / "h.bin"), "-s")
    rec = read_history(tmp_path / "h.bin", n)
    assert len(rec) == 10 * n
    spec = read_case(case)
    for r in (1, 5, 9):
        ref = oracle.run(spec, 20 * r)
        for i in range(n):
            got = rec[r * n + i]
            assert got[2] == i
            assert list(got[3:6]) == list(ref["position"][i]), (r, i)
            assert list(got[9:12]) == list(ref["velocity"][i]), (r, i)
    lag = np.array([x[17] for x in rec])            # lag_angle
    mig = np.array([x[19] for x in rec])            # migration_timescale
    body = np.array([x[2] for x in rec])
    later = np.array([x[0] for x in rec]) > 0.0
    if "disk" in name:
        assert np.all(mig[later & (body > 0)] > 0) and np.all(mig[body == 0] == 0)
    elif "bolmont" in name:
        assert np.all(lag[body == 0] > 0) and np.all(lag[body > 0] == 0)
    else:
        assert np.all(lag == 0) and np.all(mig == 0)


def test_resume_with_orbiting_dynamical_tides(tmp_path):
    """An orbiting body with LeconteChabrier2013 dynamical tides: a run stopped
    early and resumed from its snapshot (which carries the body's
    pair-dependent map entry, constant_time_lag.rs:188-194) writes the same
    history as an uninterrupted run."""
    with open(os.path.join(GOLDEN_DIR, "test_evolution-solar_like_bolmontmathis2016.json")) as f:
        case = json.load(f)["case"]
    u = case["universe"]
    u["particles"][2]["evolution"] = {"LeconteChabrier2013": True}
    ev = u["particles_evolvers"][2]
    ev["evolution"] = {"LeconteChabrier2013": True}
    ev["inverse_tidal_q_factor"] = [0.1] * len(ev["time"])
    n = u["n_particles"]
    dt = case["time_step"]
    case["historic_snapshot_period"] = 10 * dt
    case["recovery_snapshot_period"] = 50 * dt
    with open(tmp_path / "full.json", "w") as f:
        json.dump(case, f)
    cli("start", str(tmp_path / "full.json"), str(tmp_path / "a.json"), str(tmp_path / "a.bin"), "-s")
    limit = u["time_limit"]
    u["time_limit"] = limit * 0.6
    with open(tmp_path / "short.json", "w") as f:
        json.dump(case, f)
    cli("start", str(tmp_path / "short.json"), str(tmp_path / "b.json"), str(tmp_path / "b.bin"), "-s")
    snap = json.load(open(tmp_path / "b.json"))
    entries = snap["universe"]["pair_dependent_scaled_dissipation_factor"]
    assert list(entries) == ["20"]     # body id 2, host id 0
    cli("resume", str(tmp_path / "b.json"), str(tmp_path / "b.bin"), "-s", "--time-limit", repr(limit))
    assert open(tmp_path / "b.bin", "rb").read() == open(tmp_path / "a.bin", "rb").read()
    assert len(read_history(tmp_path / "a.bin", n)) > 0


def test_bincode_snapshots_resume_bit_for_bit(tmp_path):
    """The reference's documented usage, `start case.json case.bin history.bin`
    (README.md:55-58): a snapshot whose name does not end in .json is written
    in the bincode layout (output.rs:45-52), the previous one is kept as
    <stem>.<YYYYMMDD>T<AM|PM>.bin (35-41), and a run stopped early and resumed
    from it writes the same history as an uninterrupted run."""
    import glob
    full = write_case(tmp_path / "case.json", 10, 50)
    n = full["universe"]["n_particles"]
    cli("start", str(tmp_path / "case.json"), str(tmp_path / "a.bin"), str(tmp_path / "a.hist"), "-s")
    write_case(tmp_path / "short.json", 10, 50, time_limit=full["universe"]["time_limit"] * 0.6)
    cli("start", str(tmp_path / "short.json"), str(tmp_path / "b.bin"), str(tmp_path / "b.hist"), "-s")
    raw = open(tmp_path / "b.bin", "rb").read()
    with pytest.raises(ValueError):     # a .bin snapshot is not JSON
        json.loads(raw)
    assert struct.unpack_from("<d", raw, 0)[0] == full["time_step"]
    assert glob.glob(str(tmp_path / "b.*T[AP]M.bin")), "the previous snapshot is kept as a 12-hourly backup"
    cli("resume", str(tmp_path / "b.bin"), str(tmp_path / "b.hist"), "-s", "--time-limit",
        repr(full["universe"]["time_limit"]))
    assert open(tmp_path / "b.hist", "rb").read() == open(tmp_path / "a.hist", "rb").read()
    assert len(read_history(tmp_path / "a.hist", n)) > 0


def test_json_snapshot_and_its_bincode_form_resume_identically(tmp_path):
    """A JSON recovery snapshot converted to bincode (`posidonius-hip convert`)
    resumes to the same history, byte for byte, as the JSON one."""
    import shutil
    full = write_case(tmp_path / "case.json", 10, 50)
    write_case(tmp_path / "short.json", 10, 50, time_limit=full["universe"]["time_limit"] * 0.6)
    cli("start", str(tmp_path / "short.json"), str(tmp_path / "s.json"), str(tmp_path / "s.hist"), "-s")
    n = full["universe"]["n_particles"]
    snap = json.load(open(tmp_path / "s.json"))
    # the input has no ids; the snapshot holds every particle's, its index (universe.rs:119)
    assert [p["id"] for p in snap["universe"]["particles"][:n]] == list(range(n))
    cli("convert", str(tmp_path / "s.json"), str(tmp_path / "s.bin"))
    shutil.copy(tmp_path / "s.hist", tmp_path / "t.hist")
    limit = repr(full["universe"]["time_limit"])
    cli("resume", str(tmp_path / "s.json"), str(tmp_path / "s.hist"), "-s", "--time-limit", limit)
    cli("resume", str(tmp_path / "s.bin"), str(tmp_path / "t.hist"), "-s", "--time-limit", limit)
    assert open(tmp_path / "t.hist", "rb").read() == open(tmp_path / "s.hist", "rb").read()
