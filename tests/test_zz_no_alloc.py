"""pos_additional_effects_noalloc against pos_additional_effects, bit for bit,
through the C ABI: two calls in a row on one GPU (the second with a smaller
ensemble, the scratch array keeping the first call's contents), outputs in
caller-owned device buffers, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from posidonius_amd import abi, generators as gen
from posidonius_amd.case import EnsembleSpec

pytestmark = pytest.mark.gpu


def _ensemble(n, seed):
    return EnsembleSpec.concat([gen.config2(n, seed=seed), gen.config3(max(1, n // 8), seed=seed + 1),
                                gen.config1(max(1, n // 4), seed=seed + 2)])


def test_noalloc_equals_the_allocating_entry_point():
    import torch
    from posidonius_amd.ensemble import Ensemble
    lib = abi.library()
    for n, seed in ((2048, 101), (96, 102)):
        spec = _ensemble(n, seed)
        nb = spec.n_bodies
        print(f"noalloc: {spec.n_systems} systems, {nb} bodies", flush=True)
        with Ensemble(spec, mapping="packed", group_size=-1, steps_per_launch=16) as ens:
            ens.step(5)
            acc, dl = ens.additional_effects()
            out_a = torch.full((nb, 3), float("nan"), dtype=torch.float64, device="cuda")
            out_l = torch.full((nb, 3), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            rc = lib.pos_additional_effects_noalloc(ens._h, C.c_void_p(out_a.data_ptr()), C.c_void_p(out_l.data_ptr()))
            abi.check(rc)
            ens.synchronize()
            np.testing.assert_array_equal(out_a.cpu().numpy(), acc)
            np.testing.assert_array_equal(out_l.cpu().numpy(), dl)
            # the state is untouched: the steps after it equal those without it
            ens.step(5)
            after = ens.get_state(fields=("position", "velocity"))
        with Ensemble(spec, mapping="packed", group_size=-1, steps_per_launch=16) as ref:
            ref.step(10)
            ref_state = ref.get_state(fields=("position", "velocity"))
        for k in after:
            np.testing.assert_array_equal(after[k], ref_state[k])


def test_noalloc_rejects_ensembles_beyond_its_scratch():
    import torch
    from posidonius_amd.ensemble import Ensemble
    lib = abi.library()
    n_sys = abi.NOALLOC_MAX_BODIES // 3 + 1          # star + 2 planets each: one body over the limit
    spec = gen.config1(n_sys, seed=5)
    assert spec.n_bodies > abi.NOALLOC_MAX_BODIES
    out = torch.empty((1, 3), dtype=torch.float64, device="cuda")
    with Ensemble(spec, mapping="packed") as ens:
        rc = lib.pos_additional_effects_noalloc(ens._h, C.c_void_p(out.data_ptr()), C.c_void_p(out.data_ptr()))
        assert rc == abi.E_LIMIT
        assert b"scratch" in lib.pos_last_error()
