"""Write compact golden fixtures from the reference's own test data.

Reads, for every WHFast fixture set under
``/root/reference/tests/data/<test>/`` (case.json = the integrator state
written by tests/common/universe.rs:37-44, particle_<i>.json = inertial
position/velocity/acceleration after iterating to the time limit,
tests/common/universe.rs:19-34), and writes ``tests/golden/<test>.json``:

* the case in the reference's own JSON schema, reduced to the keys the
  case reader uses and to the first ``n_particles`` particles;
* evolution tables cut to a window of rows around the integration interval
  (the interpolation result depends only on the bracketing rows; the
  fixture test re-checks bit-equality of the trimmed cases);
* the expected states from particle_<i>.json.

Run from the repository root: ``python tests/golden/make_fixtures.py``.
Only data files are read; no reference code is imported or executed.
"""
import glob
import json
import os
import sys

REF = "/root/reference/tests/data"
OUT = os.path.dirname(os.path.abspath(__file__))

PARTICLE_KEYS = ("mass", "radius", "radius_of_gyration_2", "moment_of_inertia", "inertial_position",
                 "inertial_velocity", "spin", "angular_momentum", "heliocentric_velocity", "evolution")
EFFECT_KEYS = ("tides", "rotational_flattening", "general_relativity", "disk")
WINDOW = 12


def trim_evolver(ev, t_end):
    t = ev["time"]
    if not t:
        return ev
    # first row strictly after 0 and after t_end (tools.rs:882-905 search)
    lo = next((i for i, x in enumerate(t) if x > 0.0), len(t))
    hi = next((i for i, x in enumerate(t) if x > t_end), len(t))
    left_index = ev.get("left_index", 0)
    # Evolver::idx() searches from left_index - 10 (evolution.rs:449-456) and
    # left_index only grows; starting the window below both the bracketing row
    # at t = 0 and left_index, and shifting left_index by the same offset,
    # leaves every interpolated value unchanged.
    a = max(0, min(lo - 1, left_index) - WINDOW)
    b = min(len(t), hi + WINDOW)
    out = {"evolution": ev["evolution"], "left_index": left_index - a}
    for k in ("time", "radius", "radius_of_gyration_2", "love_number", "inverse_tidal_q_factor"):
        col = ev.get(k) or []
        out[k] = col[a:b] if len(col) == len(t) else []
    return out


def reduce_case(case):
    u = case["universe"]
    n = u["n_particles"]
    parts = []
    for p in u["particles"][:n]:
        q = {k: p[k] for k in PARTICLE_KEYS}
        for e in EFFECT_KEYS:
            q[e] = {"effect": p[e]["effect"]}
        q["wind"] = {"effect": p["wind"]["effect"], "parameters": {"input": p["wind"]["parameters"]["input"]}}
        parts.append(q)
    t_end = u["time_limit"]
    evolvers = [trim_evolver(ev, t_end) for ev in u["particles_evolvers"][:n]]
    keep_top = ("alternative_coordinates_type", "current_iteration", "current_time", "half_time_step",
                "historic_snapshot_period", "last_historic_snapshot_time", "recovery_snapshot_period",
                "time_step", "timestep_warning")
    out = {k: case[k] for k in keep_top}
    out["inertial_velocity_errors"] = case["inertial_velocity_errors"][:n]
    out["particle_angular_momentum_errors"] = case["particle_angular_momentum_errors"][:n]
    out["universe"] = {
        "consider_effects": u["consider_effects"],
        "general_relativity_implementation": u["general_relativity_implementation"],
        "initial_time": u["initial_time"],
        "n_particles": n,
        "time_limit": u["time_limit"],
        "particles": parts,
        "particles_evolvers": evolvers,
    }
    return out


def main():
    if not os.path.isdir(REF):
        sys.exit(f"{REF} not found")
    written = 0
    for d in sorted(glob.glob(os.path.join(REF, "*"))):
        cj = os.path.join(d, "case.json")
        if not os.path.exists(cj):
            continue
        with open(cj) as f:
            case = json.load(f)
        if "alternative_coordinates_type" not in case:
            continue   # IAS15 / LeapFrog cases: other integrators
        n = case["universe"]["n_particles"]
        expected = []
        for i in range(n):
            with open(os.path.join(d, f"particle_{i}.json")) as f:
                pj = json.load(f)
            expected.append({k: [pj[k]["x"], pj[k]["y"], pj[k]["z"]]
                             for k in ("inertial_position", "inertial_velocity", "inertial_acceleration")})
        name = os.path.basename(d)
        out = {"name": name, "source": f"tests/data/{name}", "case": reduce_case(case), "expected": expected}
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump(out, f, separators=(",", ":"))
        written += 1
    print(f"wrote {written} fixtures to {OUT}")


if __name__ == "__main__":
    main()
