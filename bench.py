#!/usr/bin/env python3
"""Benchmark of the WHFast + full-effects step on one MI355X.

Workload (default): the uneven-size ensemble of 65,536 independent systems,
each a 0.09 M_sun star with 1-11 Earth-mass planets in compact chains (2-12
bodies), with constant-time-lag tides, rotational flattening, Kidder 1995 GR
and spin integration, in FP64.  A "step" is one WHFast iterate of every
system; the metric is system-steps per second of the batched (packed,
multi-step, graph-replayed, autotuned) kernel, with the plain mapping (one
system per wavefront, one step per launch) measured beside it.

Printed: one JSON line on stdout.  Progress goes to stderr.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--systems B]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 operations (+ - * / sqrt, each counted once; nothing is
# fused) per system-step, by body count, for the uneven-size ensemble.  They
# are the means measured by the oracle's FlopCount instantiation over the
# workload's own systems (256 systems x 100 steps, seed 1002), i.e. for the
# iteration counts the data actually take; tests/test_flop_table.py re-derives
# them.
FLOPS_PER_SYSTEM_STEP = {2: 5978, 3: 11356, 4: 16747, 5: 22177, 6: 27614, 7: 33118, 8: 38616,
                         9: 44177, 10: 49766, 11: 55408, 12: 61076}
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector, AMD specification (an FMA counts 2)
# Without multiply-add contraction (the step is built with -ffp-contract=off,
# as the reference never fuses) every operation is one add or multiply, and
# the FP64 pipe retires half the flops it does on FMAs.
FP64_NO_FMA_PEAK_TFLOPS = FP64_PEAK_TFLOPS / 2
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HipEvents:
    """Two HIP events recorded on an explicit stream (the library's own)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        for e in (self.e0, self.e1):
            if self.hip.hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")

    def record(self, which, stream):
        if self.hip.hipEventRecord(self.e0 if which == 0 else self.e1, C.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self):
        self.hip.hipEventSynchronize(self.e1)
        ms = C.c_float()
        if self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)


def timed_steps(ens, steps, events):
    ens.synchronize()
    stream = ens.stream()
    t0 = time.perf_counter()
    events.record(0, stream)
    ens.step(steps)
    events.record(1, stream)
    ens.synchronize()
    wall = time.perf_counter() - t0
    return wall, events.elapsed_ms()


def measured_traffic(workload, launch_steps):
    """HBM bytes of one step-kernel launch of `launch_steps` steps, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries of the same workload
    (profiles/*.md, `traffic_json:` lines, counters already scaled by the
    calibration copy).  A launch moves a fixed amount (the state in and out,
    the set-up around the step loop) plus an amount per step, so two
    summaries of different launch lengths give both, and any launch length
    is priced from them; one summary of exactly that length is used as is.
    None when the summaries do not cover it."""
    import glob
    recs = []
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_json:"):
                    rec = json.loads(line.split(":", 1)[1])
                    if rec["workload"] == workload:
                        rec["source"] = os.path.relpath(path, ROOT)
                        recs.append(rec)
    for rec in recs:
        if rec["steps_per_launch"] == launch_steps:
            return dict(rec, model="measured at this launch length")
    lengths = sorted({r["steps_per_launch"] for r in recs})
    if len(lengths) < 2:
        return None
    a = next(r for r in recs if r["steps_per_launch"] == lengths[0])
    b = next(r for r in recs if r["steps_per_launch"] == lengths[-1])
    per_step = (b["bytes_per_launch"] - a["bytes_per_launch"]) / (lengths[-1] - lengths[0])
    fixed = a["bytes_per_launch"] - per_step * lengths[0]
    return {"workload": workload, "steps_per_launch": launch_steps, "bytes_per_launch": fixed + per_step * launch_steps,
            "fetch_scale": a.get("fetch_scale"), "write_scale": a.get("write_scale"),
            "source": f"{a['source']}, {b['source']}",
            "model": f"{fixed:.4g} B per launch + {per_step:.4g} B per step, from launches of "
                     f"{lengths[0]} and {lengths[-1]} steps"}


def adjoint_cost(n_systems, n_steps, segment_steps):
    """Cost of the discrete adjoint against the forward pass on the adjoint
    workload (star + 3 planets, DH coordinates, tides + flattening + GR):
    forward = pos_step over n_steps; recording = the same steps with a state
    checkpoint every segment_steps; backward = pos_adjoint (recomputation of
    every segment with its step tape, then the reverse sweep)."""
    import numpy as np
    from posidonius_amd import generators as gen
    from posidonius_amd.ensemble import Ensemble
    spec = gen.config4(n_systems)
    with Ensemble(spec, mapping="packed", use_graph=True) as ens:
        ens.step(1)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.step(n_steps)
        ens.synchronize()
        fwd = time.perf_counter() - t0
    with Ensemble(spec, mapping="packed", use_graph=True) as ens:
        ens.step(1)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.adjoint_record(n_steps, segment_steps)
        rec = time.perf_counter() - t0
        rng = np.random.default_rng(0)
        seed = rng.standard_normal((spec.n_bodies, 3))
        t0 = time.perf_counter()
        g = ens.adjoint(position=seed, velocity=seed, angular_momentum=None)
        bwd = time.perf_counter() - t0
    finite = bool(np.all(np.isfinite(g["position"])) and np.all(np.isfinite(g["dissipation_factor"])))
    return {"workload": f"{n_systems} star + 3 planet systems (tides, flattening, Kidder GR, spin), "
                        f"{n_steps} steps, checkpoint every {segment_steps} steps",
            "forward_s": fwd, "record_s": rec, "backward_s": bwd,
            "record_over_forward": rec / fwd, "backward_over_forward": bwd / fwd,
            "forward_system_steps_per_s": n_systems * n_steps / fwd, "gradients_finite": finite,
            "target_backward_over_forward": 3.0}


def cpu_baseline(spec, threads, target_s=12.0):
    """The oracle (C++ restatement of the reference step) on the host cores:
    one system per thread at a time, `threads` threads, a bounded sample."""
    import numpy as np
    from tests.oracle_runner import Oracle
    oracle = Oracle()
    n = threads * 16
    sample = spec.subset(np.arange(n))
    t0 = time.perf_counter()
    oracle.run(sample, 20, threads=threads, fields=("status",))
    calib = time.perf_counter() - t0
    steps = max(20, int(20 * target_s / max(calib, 1e-3)))
    t0 = time.perf_counter()
    out = oracle.run(sample, steps, threads=threads, fields=("status",))
    wall = time.perf_counter() - t0
    done = int(out["steps_done"].sum())
    # the one-system CPU case (an M dwarf with two Earth-mass planets, full
    # effects), 1e5 steps on one core
    from posidonius_amd import generators as gen
    one = gen.config0()
    t0 = time.perf_counter()
    r0 = oracle.run(one, 100_000, threads=1, fields=("status",))
    wall0 = time.perf_counter() - t0
    return {"value": done / wall, "unit": "system-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} systems of the same ensemble (first {n}), {steps} steps each, "
                      f"{threads} threads, one system per thread at a time; {wall:.1f} s wall",
            "build": "g++ -O3 -ffp-contract=off (oracle/Makefile)",
            "host": {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
                     "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None},
            "single_system": {"workload": "M dwarf + 2 Earth-mass planets, tides + flattening + Kidder GR, "
                                          "dt 0.08 d", "steps": int(r0["steps_done"][0]), "wall_s": wall0,
                              "cores": 1, "system_steps_per_s": float(r0["steps_done"][0]) / wall0}}


def operator_rates(ens, n_sys, calls=5):
    """The operator-level entry points the reference benchmarks per universe
    (benches/simulation.rs:127 gravity_calculate_acceleration, 138
    calculate_additional_effects): pos_gravity_acceleration and
    pos_additional_effects over every system of the ensemble.  The kernels
    are timed with HIP events on the library's stream (the handle's own
    pos_get_launch_stats); the whole call, which also reorders the result
    into the caller's body order and copies it to host memory, is timed on
    the host clock and reported beside it."""
    out = {}
    for name, call in (("gravity", ens.gravity_acceleration), ("additional_effects", ens.additional_effects)):
        call()
        kernel = []
        t0 = time.perf_counter()
        for _ in range(calls):
            call()
            kernel.append(ens.launch_stats()["last_operator_kernel_ms"])
        wall = (time.perf_counter() - t0) / calls
        k_ms = sum(kernel) / len(kernel)
        out[name] = {"kernel_ms": k_ms, "kernel_systems_per_s": n_sys / (k_ms * 1e-3),
                     "call_ms": wall * 1e3, "call_systems_per_s": n_sys / wall,
                     "copy_and_reorder_ms": wall * 1e3 - k_ms}
    out["note"] = ("one call evaluates every system; kernel_* is the device time of the evaluation kernels, "
                   "call_* adds the reordering into the caller's body order and the copy to pageable host memory")
    return out


GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
# The effect sets of the reference's operator benchmark group
# calculate_additional_effects (benches/simulation.rs:138-221; its Kaula
# universe is out of scope here), by the ABI's effect bits.
EFFECT_SETS = (("all", 0b111111), ("constant_time_lag_tides", 1 << 0), ("rotational_flattening", 1 << 1),
               ("general_relativity", 1 << 2), ("disk", 1 << 3), ("wind", 1 << 4), ("evolution", 1 << 5),
               ("none", 0))


def benchmark_universe():
    """The universe of the reference's operator benchmark
    (benches/simulation.rs:14-46): a solar-like star with a protoplanetary
    disk, GalletBolmont2017 evolution and Kidder 1995 GR, and the common
    four-planet configuration, at t = 0.  The committed disk fixture holds that
    star and those planets with Baraffe1998 evolution; the star's evolution
    and its evolution table are taken from the GalletBolmont2017 fixture."""
    import copy
    from posidonius_amd.case import read_case
    with open(os.path.join(GOLDEN_DIR, "test_disk-enabled_disk.json")) as f:
        case = copy.deepcopy(json.load(f)["case"])
    with open(os.path.join(GOLDEN_DIR, "test_evolution-solar_like_galletbolmont2017.json")) as f:
        gb = json.load(f)["case"]
    star, gb_star = case["universe"]["particles"][0], gb["universe"]["particles"][0]
    for key in ("evolution", "radius", "radius_of_gyration_2", "moment_of_inertia"):
        star[key] = gb_star[key]
    case["universe"]["particles_evolvers"][0] = gb["universe"]["particles_evolvers"][0]
    return read_case(case)


def effect_operator_rates(n_copies, calls=3):
    """pos_additional_effects over n_copies of the benchmark universe, once per
    effect set of the reference's benchmark group: the evaluation kernels'
    device time (HIP events on the library's stream), as kernel_ms and
    systems per second."""
    from posidonius_amd.case import EnsembleSpec
    from posidonius_amd.ensemble import Ensemble
    spec = benchmark_universe()
    while spec.n_systems < n_copies:
        spec = EnsembleSpec.concat([spec, spec])
    out = {"workload": f"{spec.n_systems} copies of the benchmark universe (star with disk + 4 planets)",
           "note": "the spin derivatives are evaluated as the step evaluates them (only when spins are "
                   "integrated: tides, flattening, evolution or Kidder GR); the reference's benchmark asks "
                   "for them in every set, which differs for the wind set alone"}
    for name, mask in EFFECT_SETS:
        s = spec.copy()
        s.system["effects"] = s.system["effects"].copy()
        s.system["effects"][:] = mask
        with Ensemble(s, mapping="packed", group_size=-1) as ens:
            ens.additional_effects()
            kernel = []
            for _ in range(calls):
                ens.additional_effects()
                kernel.append(ens.launch_stats()["last_operator_kernel_ms"])
        k_ms = sum(kernel) / len(kernel)
        out[name] = {"kernel_ms": k_ms, "kernel_systems_per_s": s.n_systems / (k_ms * 1e-3)}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4])
    ap.add_argument("--systems", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-operators", action="store_true")
    ap.add_argument("--adjoint-systems", type=int, default=1024)
    ap.add_argument("--adjoint-steps", type=int, default=100000)
    ap.add_argument("--adjoint-segment", type=int, default=1000)
    ap.add_argument("--group-size", type=int, default=0)
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true", help="launch the K-step kernels directly, without a hipGraph")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB mebibytes five times "
                         "(reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.randn(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                          "launches": 5}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: every system is independent, so the ensemble "
              "shards by systems with no exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from posidonius_amd import abi, generators as gen
    from posidonius_amd.ensemble import Ensemble

    defaults = {1: 4096, 2: 65536, 3: 16384, 4: 1024}
    n_sys = args.systems or defaults[args.config]
    log(f"generating config {args.config} with {n_sys} systems")
    spec = gen.CONFIGS[args.config](n_sys)
    counts = spec.n_per_system()
    flops_per_step = float(sum(FLOPS_PER_SYSTEM_STEP[int(n)] for n in counts))
    events = HipEvents()

    # ---- batched path -----------------------------------------------------
    log("creating the batched ensemble")
    ens = Ensemble(spec, mapping="packed", group_size=args.group_size,
                   steps_per_launch=args.steps_per_launch, use_graph=not args.no_graph)
    if not args.no_tune:
        log("autotuning the launch mapping")
        o = ens.tune(probe_steps=32)
        log(f"chosen group_size={o.group_size} steps_per_launch={o.steps_per_launch}")
    o = ens.options()
    K = o.steps_per_launch
    log(f"warmup {args.warmup} steps")
    ens.step(args.warmup)
    ens.synchronize()
    log(f"timing {args.steps} steps")
    before = ens.launch_stats()
    wall, ev_ms = timed_steps(ens, args.steps, events)
    after = ens.launch_stats()
    direct = int(after["kernel_launches"] - before["kernel_launches"])
    replays = int(after["graph_replays"] - before["graph_replays"])
    launches = direct + replays * int(after["launches_per_replay"])
    t0 = time.perf_counter()
    full_state = ens.get_state()     # every field, device -> host
    readback_s = time.perf_counter() - t0
    frozen = int(np.count_nonzero(full_state["status"] & abi.STATUS_FATAL))
    del full_state
    operators = None
    if not args.no_operators:
        operators = operator_rates(ens, n_sys)
        log(f"operators (kernels): gravity {operators['gravity']['kernel_systems_per_s']:.4g} systems/s, "
            f"additional effects {operators['additional_effects']['kernel_systems_per_s']:.4g} systems/s")
    ens.close()
    if operators is not None:
        log("additional effects per effect set, on the reference benchmark's universe")
        operators["per_effect"] = effect_operator_rates(65536)
        log(", ".join(f"{k} {v['kernel_ms']:.3f} ms" for k, v in operators["per_effect"].items() if isinstance(v, dict)))
    value = n_sys * args.steps / wall
    # the dominant kernel: one launch per variant present (one here), each of
    # launch_steps steps; the average launch is the timed region / launches
    launch_steps = min(K, args.steps)
    kernel_ms = ev_ms / max(1, launches)
    flops_per_launch = flops_per_step * launch_steps
    achieved_tflops = flops_per_launch / (kernel_ms * 1e-3) / 1e12
    log(f"batched: {value:.4g} system-steps/s, {wall / args.steps * 1e3:.3f} ms/step, "
        f"kernel {kernel_ms:.3f} ms/launch, {achieved_tflops:.2f} TFLOP/s algorithmic")

    # ---- plain path ---------------------------------------------------------
    plain = None
    if not args.no_plain:
        log("creating the plain ensemble (one system per wavefront, one step per launch)")
        pens = Ensemble(spec, mapping="plain")
        pens.step(1)
        pwall, pms = timed_steps(pens, args.plain_steps, events)
        pens.close()
        prate = n_sys * args.plain_steps / pwall
        plain = {"value": prate, "unit": "system-steps/s", "ms_per_step": pwall / args.plain_steps * 1e3,
                 "steps": args.plain_steps, "mapping": "one system per wavefront, one step per launch, "
                                                      "state in HBM between steps"}
        log(f"plain: {prate:.4g} system-steps/s ({pwall / args.plain_steps * 1e3:.2f} ms/step)")

    # ---- differentiable form: adjoint / forward cost ---------------------------
    adj = None
    if not args.no_adjoint:
        log(f"adjoint cost: {args.adjoint_systems} systems x {args.adjoint_steps} steps")
        adj = adjoint_cost(args.adjoint_systems, args.adjoint_steps, args.adjoint_segment)
        log(f"adjoint: forward {adj['forward_s']:.2f} s, record {adj['record_s']:.2f} s, "
            f"backward {adj['backward_s']:.2f} s ({adj['backward_over_forward']:.2f}x forward)")

    # ---- CPU baseline ---------------------------------------------------------
    cpu = None
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        log(f"CPU baseline on {threads} host threads")
        cpu = cpu_baseline(spec, threads)
        log(f"cpu: {cpu['value']:.4g} system-steps/s")

    mean_n = float(counts.mean())
    workload_key = {1: "uniform", 2: "uneven", 3: "binary", 4: "adjoint"}[args.config] + str(n_sys)
    traffic = measured_traffic(workload_key, launch_steps) if args.steps % launch_steps == 0 else None
    result = {
        "metric": "ensemble system-steps/s FP64 full effects (batched vs plain GPU); adjoint/fwd cost",
        "value": value,
        "unit": "system-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": {1: "uniform ensemble: 4096 star + 2 planet systems",
                         2: "uneven-size ensemble: 65536 systems of 2-12 bodies (1 star + 1-11 planets)",
                         3: "binary-host ensemble: 16384 systems, 2 stars + 1-3 circumbinary planets",
                         4: "adjoint workload forward pass: 1024 star + 3 planet systems"}[args.config]
                        if n_sys == defaults[args.config] else f"config {args.config} with {n_sys} systems",
            "systems": n_sys, "mean_bodies": mean_n, "effects": "CTL tides + oblate flattening + Kidder GR + spin",
            "coordinates": "Jacobi" if args.config == 3 else "DemocraticHeliocentric",
            "group_size": o.group_size, "steps_per_launch": K,
            "timed_region": {"steps": args.steps, "kernel_launches_direct": direct, "graph_replays": replays,
                             "launches_per_replay": int(after["launches_per_replay"]) if replays else 0,
                             "steps_per_kernel_launch": launch_steps},
            "hip_graph": replays > 0,
        },
        "roofline": {
            "bound": "fp64-valu",
            "note": "FP64 vector-ALU bound: nothing in the step is a contraction, no MFMA is used. peak is the "
                    "MI355X FP64 vector peak (an FMA counts 2); peak_no_fma is the ceiling of separately rounded "
                    "adds and multiplies, which is what the step executes (-ffp-contract=off, as the reference)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "peak_no_fma": FP64_NO_FMA_PEAK_TFLOPS,
            "frac_no_fma": achieved_tflops / FP64_NO_FMA_PEAK_TFLOPS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_source": (f"rocprofv3 FETCH_SIZE x{traffic.get('fetch_scale', 2.0)} + WRITE_SIZE x"
                               f"{traffic.get('write_scale', 1.0)} per launch of {launch_steps} steps, "
                               f"{traffic['model']} ({traffic['source']})") if traffic else None,
            "achieved_hbm_gbs": (traffic["bytes_per_launch"] / (kernel_ms * 1e-3) / 1e9) if traffic else None,
            "hbm_peak_gbs": HBM_PEAK_GBS,
            "kernel": "whfast_steps_kernel",
            "kernel_ms_per_launch": kernel_ms,
            "steps_per_launch": launch_steps,
            "flops_per_launch": flops_per_launch,
            "flops_per_system_step_mean": flops_per_step / n_sys,
        },
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "adjoint": adj,
        "operators": operators,
        "frozen_systems": frozen,
        "pcie_inclusive": {"value": n_sys * args.steps / (wall + readback_s), "unit": "system-steps/s",
                           "readback_ms": readback_s * 1e3,
                           "note": "the timed steps plus one pos_get_state of every field to host memory"},
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
