"""Host broadphase tree: the deferred enlargement used by the refit (EnlargeProxyDeferred) gives
the same tree as immediate enlargement in the reference's form (src/dynamic_tree.c, EnlargeProxy),
whatever the enlargement order. Compiled from box2d_amd/csrc/tree.cpp with g++ and run on the
CPU."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("n", [50, 2000, 12000])
def test_deferred_enlargement_builds_the_same_tree(tmp_path, n):
    exe = str(tmp_path / "tree_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "box2d_amd", "csrc"), os.path.join(ROOT, "tests", "host", "tree_check.cpp"),
                    os.path.join(ROOT, "box2d_amd", "csrc", "tree.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe, str(n)], check=True, capture_output=True, text=True).stdout.split()
    assert len(out) == 3
    assert out[0] == out[1] == out[2], f"immediate {out[0]}, deferred {out[1]}, deferred reversed {out[2]}"
