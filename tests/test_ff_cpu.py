"""CPU: the exact fast-forward of the edge recurrence e += n.x.

mr::ff_add (minirender_amd/csrc/exact.h) replaces n dependent fp32 adds by
binade-wise jumps.  A g++ build of the same header must agree bit for bit
with the serial loop on random, tie-heavy and special-value cases.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def ff_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("ff") / "ff_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    f"-I{ROOT / 'minirender_amd' / 'csrc'}", str(ROOT / "tests" / "cpu" / "ff_check.cpp"),
                    "-o", str(out)], check=True)
    return out


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_ff_matches_serial(ff_bin, seed):
    r = subprocess.run([str(ff_bin), "150000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
