"""CPU-side checks of the C ABI: the library loads, exports every function
declared in include/posidonius_hip.h, and reports its ABI version.  No call
here touches a GPU kernel."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from posidonius_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "posidonius_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pos_[a-z_]+)\s*\(", text)))


def test_header_declares_the_expected_entry_points():
    assert set(declared_functions()) == set(abi.EXPORTED_SYMBOLS)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        pytest.fail(f"{abi.LIB_PATH} not built; run __graft_entry__.build()")
    return abi.library()


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_abi_version(lib):
    assert lib.pos_abi_version() == abi.ABI_VERSION


def test_ctypes_struct_layout_matches_header():
    # pointer fields only after the leading int32 -> 8-byte aligned
    n_ptr_desc = len(abi.EnsembleDesc._fields_) - 1
    assert C.sizeof(abi.EnsembleDesc) == 8 + 8 * n_ptr_desc
    assert C.sizeof(abi.State) == 8 * len(abi.State._fields_)
    assert C.sizeof(abi.Options) == 8 * 4
    # field order against the header text
    text = open(HEADER).read()
    desc_block = text[text.index("typedef struct pos_ensemble_desc"):text.index("} pos_ensemble_desc;")]
    names = re.findall(r"\*\s*([a-z_0-9]+)\s*;", desc_block)
    assert names == [f for f, _ in abi.EnsembleDesc._fields_[1:]]
    state_block = text[text.index("typedef struct pos_state"):text.index("} pos_state;")]
    names = re.findall(r"\*\s*([a-z_0-9]+)\s*;", state_block)
    assert names == [f for f, _ in abi.State._fields_]


def test_null_arguments_are_rejected_without_a_device(lib):
    out = C.c_void_p()
    assert lib.pos_create(None, None, C.byref(out)) == abi.E_INVALID
    assert b"NULL" in lib.pos_last_error()
    assert lib.pos_step(None, 1) == abi.E_INVALID


def test_invalid_desc_is_rejected_before_any_device_work(lib):
    from posidonius_amd import generators as gen
    spec = gen.config1(4)
    spec.system["effects"] = spec.system["effects"] | abi.EFFECT_DISK
    h = spec.desc()
    out = C.c_void_p()
    rc = lib.pos_create(h.ref, None, C.byref(out))
    assert rc == abi.E_INVALID
    assert b"disk_properties" in lib.pos_last_error()
    # properties given, but the host is no disk central body
    spec.system["disk_properties"] = np.tile([0.01, 100.0, 3.6525e7, 0.01, 1e-4, 2.4], (spec.n_systems, 1))
    h = spec.desc()
    assert lib.pos_create(h.ref, None, C.byref(out)) == abi.E_UNSUPPORTED
    assert b"disk host" in lib.pos_last_error()


def test_evolving_dissipation_needs_its_q_factor_table(lib):
    from tests.oracle_runner import load_golden
    spec, _ = load_golden("test_evolution-solar_like_bolmontmathis2016")
    spec.body["evolution_type"][2] = abi.EVO_LECONTE_CHABRIER2013_DYN   # an orbiting body: accepted
    h = spec.desc()
    h.desc.evolver_inverse_tidal_q_factor = abi.ptr(None, C.c_double)
    out = C.c_void_p()
    assert lib.pos_create(h.ref, None, C.byref(out)) == abi.E_INVALID
    assert b"evolver_inverse_tidal_q_factor" in lib.pos_last_error()
