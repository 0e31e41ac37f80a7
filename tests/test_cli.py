"""The `posidonius-hip start|resume` host: argument handling that needs no GPU
(main.rs:10-150: usage, refusal to overwrite, missing input)."""
import os
import subprocess

import pytest

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "posidonius_amd", "posidonius-hip")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_usage_without_arguments():
    r = run()
    assert r.returncode == 2 and "posidonius-hip start" in r.stderr


def test_start_refuses_existing_outputs(tmp_path):
    case = tmp_path / "case.json"
    case.write_text("{}")
    snap = tmp_path / "snap.json"
    snap.write_text("x")
    r = run("start", str(case), str(snap), str(tmp_path / "h.bin"))
    assert r.returncode == 1 and "already exists" in r.stderr


def test_start_missing_case(tmp_path):
    r = run("start", str(tmp_path / "nope.json"), str(tmp_path / "s.json"), str(tmp_path / "h.bin"))
    assert r.returncode == 1 and "not possible to start" in r.stderr


def test_resume_rejects_a_document_that_is_not_a_case(tmp_path):
    snap = tmp_path / "snap.json"
    snap.write_text('{"universe": {}}')
    r = run("resume", str(snap), str(tmp_path / "h.bin"))
    assert r.returncode == 1 and "not a WHFast case" in r.stderr
