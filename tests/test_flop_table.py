"""The algorithmic-work table bench.py prices its roofline with.

bench.FLOPS_PER_SYSTEM_STEP holds, per body count, the mean number of FP64
operations one WHFast iterate of a system of the uneven-size ensemble
performs (counted by the oracle's FlopCount scalar, which counts + - * / and
sqrt once each).  This re-derives the table from the workload's own systems.
"""
import numpy as np

import bench
from posidonius_amd import generators as gen


def test_flop_table_matches_oracle_count(oracle):
    spec = gen.config2(256)
    flops, steps = oracle.flops(spec, 100)
    assert np.all(steps == 100)
    counts = spec.n_per_system()
    per_step = flops / steps
    for n in sorted(set(int(c) for c in counts)):
        mean = per_step[counts == n].mean()
        assert abs(mean - bench.FLOPS_PER_SYSTEM_STEP[n]) <= 0.005 * mean, (n, mean, bench.FLOPS_PER_SYSTEM_STEP[n])


def test_flop_count_grows_per_step(oracle):
    """Effects off: the count grows linearly with the
    number of steps after the first (every step does the same operations
    when the Kepler solver takes the same number of iterations)."""
    spec = gen.config1(4)
    spec.system["effects"][:] = 0
    f10, _ = oracle.flops(spec, 10)
    f20, _ = oracle.flops(spec, 20)
    f30, _ = oracle.flops(spec, 30)
    # the per-step count of a pure point-mass step is a few thousand operations
    per_step = (f30 - f10) / 20
    assert np.all(per_step > 500) and np.all(per_step < 20000)
    assert np.all(np.abs((f30 - f20) - (f20 - f10)) <= 0.05 * (f20 - f10))
