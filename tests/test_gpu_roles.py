"""Per-body effect roles on the device against the oracle, bit for bit.

Systems in which some planets have tides, flattening or general relativity
Disabled (the Python builder's per-body default, posidonius/particles/
particle.py:52-56; the reference's own cases/example_circumbinary_planet.py
builds such a planet) run among systems with every role set, under the
launch mappings that group bodies differently.  A planet whose tides are
Disabled while tides are on makes the host's tidal torque NaN in the
reference (its torque loop reads that planet's never-updated zero internals,
tides/common.rs:244-255); the device reproduces that too, and the system
then stops on the capped Kepler bisection, as the oracle's does.
"""
import os

import numpy as np
import pytest

from posidonius_amd import abi, generators as gen
from posidonius_amd.case import EnsembleSpec
from tests.oracle_runner import load_golden
from tests.test_gpu_parity import assert_same_state

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def Ensemble():
    from posidonius_amd.ensemble import Ensemble
    return Ensemble


def cleared(spec, body, role):
    spec = spec.copy()
    spec.body["roles"][body] &= ~np.uint32(role)
    return spec


def role_ensemble(with_nan=True):
    fx, _ = load_golden("test_tides-enabled_tides")   # star + 4 planets, all effects, evolution
    gr, _ = load_golden("test_general_relativity-kidder1995")
    parts = [
        fx,
        cleared(fx, 2, abi.ROLE_FLAT_ORBITING),
        cleared(fx, 4, abi.ROLE_GR_ORBITING),
        cleared(cleared(gr, 1, abi.ROLE_FLAT_ORBITING), 3, abi.ROLE_GR_ORBITING),
        cleared(gen.config2(6, seed=31, n_min=3, n_max=7), 1, abi.ROLE_GR_ORBITING),
        gen.config2(6, seed=32, n_min=3, n_max=7),
    ]
    if with_nan:
        parts.append(cleared(fx, 3, abi.ROLE_TIDES_ORBITING))
    # planets of the synthetic systems without flattening or GR
    c1 = gen.config1(8, seed=33)
    for s in range(0, 8, 2):
        c1.body["roles"][c1.body_offset[s] + 1] &= ~np.uint32(abi.ROLE_FLAT_ORBITING | abi.ROLE_GR_ORBITING)
    parts.append(c1)
    return EnsembleSpec.concat(parts)


@pytest.mark.parametrize("group_size,K,graph", [(-1, 64, True), (0, 16, False), (-2, 32, True)])
def test_roles_steps_bitwise(Ensemble, oracle, group_size, K, graph):
    spec = role_ensemble()
    steps = 300
    print(f"roles ensemble: {spec.n_systems} systems, G={group_size} K={K} graph={graph}, {steps} steps", flush=True)
    ref = oracle.run(spec, steps, threads=THREADS)
    with Ensemble(spec, mapping="packed", group_size=group_size, steps_per_launch=K, use_graph=graph) as ens:
        ens.step(steps)
        out = ens.get_state()
    assert_same_state(out, ref, where=f"roles G={group_size}")
    nan_sys = spec.n_systems - 9   # the tides-disabled fixture copy
    assert out["status"][nan_sys] & abi.STATUS_FATAL
    assert np.all(out["status"][:nan_sys] & abi.STATUS_FATAL == 0)


def test_roles_operators_bitwise(Ensemble, oracle):
    spec = role_ensemble()
    with Ensemble(spec, mapping="packed", group_size=-1) as ens:
        e_acc, e_dl = ens.additional_effects()
    r_acc, r_dl = oracle.additional_effects(spec)
    np.testing.assert_array_equal(e_acc, r_acc)
    np.testing.assert_array_equal(e_dl, r_dl)   # NaN where the oracle has NaN
    assert np.isnan(e_dl).any()
