"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code (CPU only).

tests/cpu/Makefile builds, with g++ and ``-fsanitize=address,undefined
-fno-sanitize-recover=all``:

* the oracle (oracle/oracle_capi.cpp) and the adjoint-math C API
  (tests/cpu/adjoint_math_capi.cpp); the oracle's and the adjoint math's own
  tests then run against those builds in a child interpreter with the
  sanitizer runtimes preloaded, so an out-of-bounds access, a use after free
  or undefined behaviour aborts the child and fails here;
* fuzz_host, a bounded mutation fuzzer of the command line's JSON parser,
  case reader and bincode snapshot decoder (tests/cpu/fuzz_host.cpp), seeded
  with reference cases: every mutant must end in a result or an exception.

No GPU code is built with a sanitizer.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SEEDS = ["test_integrator-whfast_jacobi", "test_disk-enabled_disk", "test_general_relativity-newhall1983",
         "test_evolution-solar_like_bolmontmathis2016"]
# the tests that drive the sanitized builds
DRIVEN = ["tests/test_adjoint_math.py", "tests/test_oracle_fixtures.py", "tests/test_oracle_kepler.py",
          "tests/test_oracle_roles.py", "tests/test_oracle_tangent.py", "tests/test_oracle_disk.py",
          "tests/test_oracle_dynamical_tide.py", "tests/test_flop_table.py"]


def runtime(name):
    path = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san(tmp_path_factory):
    if not (runtime("libasan.so") and runtime("libubsan.so")):
        pytest.skip("g++ sanitizer runtimes not installed")
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpu"), "san", f"OUT={out}", "-j4"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_oracle_and_adjoint_math_under_sanitizers(san):
    env = dict(os.environ)
    env.update({
        "POSIDONIUS_ORACLE_LIBRARY": str(san / "liboracle_san.so"),
        "POSIDONIUS_ADJCHECK_LIBRARY": str(san / "libadjcheck_san.so"),
        "LD_PRELOAD": f"{runtime('libasan.so')} {runtime('libubsan.so')}",
        # the interpreter itself is not instrumented: its allocations are
        # not leak-checked
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", *DRIVEN, "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in tail and "runtime error" not in tail, tail


def test_fuzz_host_input_handling(san, tmp_path):
    seeds = []
    for name in SEEDS:
        path = tmp_path / f"{name}.json"
        with open(os.path.join(GOLDEN, name + ".json")) as f:
            path.write_text(json.dumps(json.load(f)["case"]))
        seeds.append(str(path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(san / "fuzz_host"), "1500", *seeds], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    counts = [int(w) for w in r.stdout.split() if w.isdigit()]
    parsed, cases, decoded = counts
    # the mutants reach every stage: some still parse, some still read as cases
    assert parsed > 100 and cases > 10 and decoded > 100, r.stdout
