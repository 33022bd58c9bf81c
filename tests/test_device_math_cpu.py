"""CPU: device arithmetic and index helpers whose form differs from the
reference's, built with g++ from the same headers the HIP kernels use.

  ff_check       mr::ff_add / mr::advance (exact.h) replace n dependent fp32
                 adds of the edge walk e += n.x (Renderer.cpp:243) by
                 binade-wise jumps; must match the serial loop bit for bit on
                 random, tie-heavy and special-value cases.
  helpers_check  xcd_remap (common.h) is a bijection; the branch-free near
                 clip equals the reference's rotate-loop clipTriangle
                 (Renderer.cpp:131-161) bit for bit, NaNs and ties included;
                 the bbox end bound last_center equals the last pixel of the
                 reference's fp32 loop x <= pmax + 0.5f (Renderer.cpp:238,243)
                 near every power of two from 2 to 16384.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path_factory, name):
    out = tmp_path_factory.mktemp(name) / name
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    f"-I{ROOT / 'minirender_amd' / 'csrc'}", str(ROOT / "tests" / "cpu" / f"{name}.cpp"),
                    "-o", str(out)], check=True)
    return out


@pytest.fixture(scope="module")
def ff_bin(tmp_path_factory):
    return _build(tmp_path_factory, "ff_check")


@pytest.fixture(scope="module")
def helpers_bin(tmp_path_factory):
    return _build(tmp_path_factory, "helpers_check")


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_ff_matches_serial(ff_bin, seed):
    r = subprocess.run([str(ff_bin), "150000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")


def test_remap_and_clip(helpers_bin):
    r = subprocess.run([str(helpers_bin), "300000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
