"""The recovery-snapshot codec of the command-line host (host/bincode.hpp),
without a GPU.

* Its schema lists the reference's serde structs in declaration order and
  its enums' variants in index order; both are checked against
  tests/golden/serde_order.json, which a script extracts mechanically from the
  reference's Rust sources, so a field or variant out of place fails here.
* `posidonius-hip convert` takes a case through bincode and back: every value
  of the document comes back unchanged (doubles bit for bit), a second pass
  gives the same bytes, and absent fields take the reference Python writer's
  defaults (dummy particles up to MAX_PARTICLES).
There is no snapshot written by the reference binary in the reference's
tree, so byte equality with the reference itself is not pinned.
"""
import json
import os
import re
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "posidonius_amd", "posidonius-hip")
HEADER = os.path.join(ROOT, "posidonius_amd", "host", "bincode.hpp")
ORDER = os.path.join(ROOT, "tests", "golden", "serde_order.json")

# the codec's short names for the reference's struct names
RUST_NAME = {
    "TidesParameters": "TidesParticleParameters", "TidesInternal": "TidesParticleInternalParameters",
    "Output2": "TidesParticleOutputParameters", "Coordinates": "TidesParticleCoordinates",
    "FlatteningParameters": "RotationalFlatteningParticleParameters",
    "FlatteningInternal": "RotationalFlatteningParticleInternalParameters",
    "GRParameters": "GeneralRelativityParticleParameters", "GRInternal": "GeneralRelativityParticleInternalParameters",
    "WindParameters": "WindParticleParameters", "WindInput": "WindParticleInputParameters",
    "WindInternal": "WindParticleInternalParameters", "WindOutput": "WindParticleOutputParameters",
    "DiskParameters": "DiskParticleParameters", "DiskInternal": "DiskParticleInternalParameters",
    "DiskOutput": "DiskParticleOutputParameters",
}
# structs the codec shares between effects whose reference structs have the same fields
SHARED = {"Output2": ("RotationalFlatteningParticleOutputParameters", "GeneralRelativityParticleOutputParameters"),
          "Coordinates": ("RotationalFlatteningParticleCoordinates", "GeneralRelativityParticleCoordinates",
                          "DiskParticleCoordinates")}


def codec_schema():
    text = open(HEADER).read()
    out = {"S": {}, "E": {}}
    for kind, name, body in re.findall(r'\b([SE])\("(\w+)",\s*\{(.*?)\}\);', text, flags=re.S):
        out[kind][name] = re.findall(r'\{"(\w+)",\s*"[^"]*"\}', body)
    return out


def test_schema_matches_the_reference_declarations():
    ref = json.load(open(ORDER))
    mine = codec_schema()
    assert len(mine["S"]) >= 30 and len(mine["E"]) >= 11
    for name, fields in mine["S"].items():
        for rust in (RUST_NAME.get(name, name),) + SHARED.get(name, ()):
            assert ref["structs"][rust] == fields, (name, rust)
    for name, variants in mine["E"].items():
        assert ref["enums"][name] == variants, name


pytestmark_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def convert(src, dst):
    r = subprocess.run([CLI, "convert", str(src), str(dst)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def leaves(v, path=""):
    if isinstance(v, dict):
        for k, x in v.items():
            yield from leaves(x, f"{path}.{k}")
    elif isinstance(v, list):
        for i, x in enumerate(v):
            yield from leaves(x, f"{path}[{i}]")
    else:
        yield path, v


@pytestmark_cli
@pytest.mark.parametrize("name", ["test_integrator-whfast_jacobi", "test_evolution-solar_like_bolmontmathis2016",
                                  "test_disk-enabled_disk", "test_general_relativity-newhall1983"])
def test_case_round_trips_through_bincode(tmp_path, name):
    case = json.load(open(os.path.join(ROOT, "tests", "golden", name + ".json")))["case"]
    src = tmp_path / "case.json"
    src.write_text(json.dumps(case))
    convert(src, tmp_path / "case.bin")
    convert(tmp_path / "case.bin", tmp_path / "back.json")
    back = json.load(open(tmp_path / "back.json"))
    got = dict(leaves(back))
    for path, value in leaves(case):
        assert path in got, path
        if isinstance(value, float):
            assert struct.pack("<d", value) == struct.pack("<d", float(got[path])), path
        else:
            assert got[path] == value, path
    # absent fields took the defaults, and the arrays were padded to MAX_PARTICLES
    assert len(back["universe"]["particles"]) == 10 and len(back["universe"]["particles_evolvers"]) == 10
    assert back["universe"]["particles"][9]["tides"]["effect"] == "Disabled"
    assert back["universe"]["particles"][9]["evolution"] == "NonEvolving"
    assert back["universe"]["particles"][9]["wind"]["effect"] == "Disabled"
    # the first fields of the layout: time_step, half_time_step (f64), then
    # the universe's initial_time and time_limit
    raw = open(tmp_path / "case.bin", "rb").read()
    t, h, t0, tl = struct.unpack_from("<4d", raw, 0)
    assert (t, h, t0, tl) == (case["time_step"], case["half_time_step"], case["universe"]["initial_time"],
                              case["universe"]["time_limit"])
    # a second pass gives the same bytes
    convert(tmp_path / "back.json", tmp_path / "again.bin")
    assert open(tmp_path / "again.bin", "rb").read() == raw


@pytestmark_cli
def test_non_finite_numbers_are_null_in_json(tmp_path):
    case = json.load(open(os.path.join(ROOT, "tests", "golden", "test_integrator-whfast_jacobi.json")))["case"]
    case["universe"]["particles"][1]["radius"] = None   # serde_json's NaN
    src = tmp_path / "case.json"
    src.write_text(json.dumps(case))
    convert(src, tmp_path / "c.bin")
    raw = open(tmp_path / "c.bin", "rb").read()
    convert(tmp_path / "c.bin", tmp_path / "c.json")
    text = open(tmp_path / "c.json").read()
    assert "nan" not in text.lower().replace("nonevolving", "")
    assert json.loads(text)["universe"]["particles"][1]["radius"] is None
    assert raw.count(struct.pack("<d", float("nan"))) >= 1


@pytestmark_cli
def test_truncated_snapshot_is_rejected(tmp_path):
    case = json.load(open(os.path.join(ROOT, "tests", "golden", "test_integrator-whfast_jacobi.json")))["case"]
    src = tmp_path / "case.json"
    src.write_text(json.dumps(case))
    convert(src, tmp_path / "c.bin")
    raw = open(tmp_path / "c.bin", "rb").read()
    (tmp_path / "t.bin").write_bytes(raw[:len(raw) // 2])
    r = subprocess.run([CLI, "convert", str(tmp_path / "t.bin"), str(tmp_path / "t.json")], capture_output=True, text=True)
    assert r.returncode == 1 and "truncated" in r.stderr
