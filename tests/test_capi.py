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
    # properties given, but the disk's central body is a planet
    spec.system["disk_properties"] = np.tile([0.01, 100.0, 3.6525e7, 0.01, 1e-4, 2.4], (spec.n_systems, 1))
    planet = spec.body_offset[2] + 1
    spec.body["roles"][planet] |= abi.ROLE_DISK_CENTRAL
    h = spec.desc()
    assert lib.pos_create(h.ref, None, C.byref(out)) == abi.E_UNSUPPORTED
    assert b"disk host" in lib.pos_last_error()


def test_central_body_rules_of_the_reference(lib):
    """find_indices / disable_unnecessary_effects (universe.rs:699-1006): one
    central body per effect, the GR host is the most massive body, tides and
    flattening share their host; the reference panics on each of these."""
    from posidonius_amd import generators as gen
    out = C.c_void_p()
    for role, text in ((abi.ROLE_TIDES_CENTRAL, b"only one central body"),
                       (abi.ROLE_GR_CENTRAL, b"only one central body")):
        spec = gen.config1(4)
        spec.body["roles"][spec.body_offset[1] + 2] |= role
        assert lib.pos_create(spec.desc().ref, None, C.byref(out)) == abi.E_INVALID
        assert text in lib.pos_last_error()
    spec = gen.config1(4)   # GR central body moved to a planet
    b0 = spec.body_offset[3]
    spec.body["roles"][b0] &= ~np.uint32(abi.ROLE_GR_CENTRAL)
    spec.body["roles"][b0 + 1] = (spec.body["roles"][b0 + 1] & ~np.uint32(abi.ROLE_GR_ORBITING)) | abi.ROLE_GR_CENTRAL
    assert lib.pos_create(spec.desc().ref, None, C.byref(out)) == abi.E_INVALID
    assert b"most massive" in lib.pos_last_error()
    spec = gen.config1(4)   # flattening central body on a planet, tides on the star
    b0 = spec.body_offset[0]
    spec.body["roles"][b0] &= ~np.uint32(abi.ROLE_FLAT_CENTRAL)
    spec.body["roles"][b0 + 1] = (spec.body["roles"][b0 + 1] & ~np.uint32(abi.ROLE_FLAT_ORBITING)) | abi.ROLE_FLAT_CENTRAL
    assert lib.pos_create(spec.desc().ref, None, C.byref(out)) == abi.E_INVALID
    assert b"tidal & rotational flattening" in lib.pos_last_error()


def test_disabled_orbiting_roles_pass_validation(lib):
    """A body whose tides, flattening or GR is Disabled (the Python builder's
    per-body default, posidonius/particles/particle.py:52-56) is accepted:
    without a GPU the call gets past validation to the device check."""
    from posidonius_amd import generators as gen
    spec = gen.config1(4)
    for s, role in enumerate((abi.ROLE_TIDES_ORBITING, abi.ROLE_FLAT_ORBITING, abi.ROLE_GR_ORBITING)):
        spec.body["roles"][spec.body_offset[s] + 1] &= ~np.uint32(role)
    out = C.c_void_p()
    rc = lib.pos_create(spec.desc().ref, None, C.byref(out))
    assert rc in (abi.E_NO_DEVICE, abi.OK), lib.pos_last_error()
    if rc == abi.OK:
        lib.pos_destroy(out)


def test_evolving_dissipation_needs_its_q_factor_table(lib):
    from tests.oracle_runner import load_golden
    spec, _ = load_golden("test_evolution-solar_like_bolmontmathis2016")
    spec.body["evolution_type"][2] = abi.EVO_LECONTE_CHABRIER2013_DYN   # an orbiting body: accepted
    h = spec.desc()
    h.desc.evolver_inverse_tidal_q_factor = abi.ptr(None, C.c_double)
    out = C.c_void_p()
    assert lib.pos_create(h.ref, None, C.byref(out)) == abi.E_INVALID
    assert b"evolver_inverse_tidal_q_factor" in lib.pos_last_error()
