"""The minirender C++ API (include/minirender/*.h) driven as a reference user
would: tests/cpp/api_check.cpp and tests/cpp/ref_idiom.cpp.  CPU mode checks
host logic (primitive order, scene flattening, matrices, PPM, the container
idioms of the reference's samples); GPU mode renders scenes built with the
API through Renderer::render() and compares with the oracle."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BUILD = ROOT / "tests" / "cpp" / "_build"
BIN = BUILD / "api_check"


def _binary(name="api_check"):
    path = BUILD / name
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True)
    return path


def test_api_host_logic():
    r = subprocess.run([str(_binary()), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.gpu
def test_api_render_matches_oracle():
    r = subprocess.run([str(_binary()), "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_reference_idiom_host():
    """`Shared<T> x = new T()`, `<<`, shared index arrays, dup/clone, one
    object wrapped twice (samples/bench.cpp:17-48, 98)."""
    r = subprocess.run([str(_binary("ref_idiom")), "cpu"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.gpu
def test_reference_idiom_render_matches_oracle():
    r = subprocess.run([str(_binary("ref_idiom")), "gpu"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
