"""The minirender C++ API (include/minirender/*.h) driven as a reference user
would: tests/cpp/api_check.cpp.  CPU mode checks host logic (primitive
order, scene flattening, matrices, PPM); GPU mode renders scenes built with
the API through Renderer::render() and compares with the oracle."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "_build" / "api_check"


def _binary():
    if not BIN.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True)
    return BIN


def test_api_host_logic():
    r = subprocess.run([str(_binary()), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.gpu
def test_api_render_matches_oracle():
    r = subprocess.run([str(_binary()), "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
