"""Host broadphase tree: the deferred enlargement used by the refit (EnlargeProxyDeferred) gives
the same tree and the same query visit order as immediate enlargement in the reference's form
(src/dynamic_tree.c, EnlargeProxy), whatever the enlargement order. Compiled from
box2d_amd/csrc/tree.cpp with g++ and run on the CPU, once optimized and once under
AddressSanitizer and UndefinedBehaviorSanitizer (host code only; any report fails the run)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

BUILDS = {
    "O2": ["-O2"],
    "asan_ubsan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
}


@pytest.fixture(scope="module", params=sorted(BUILDS))
def tree_check(request, tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp(request.param) / "tree_check")
    subprocess.run(["g++", "-std=c++17", "-ffp-contract=off"] + BUILDS[request.param] +
                   ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "box2d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "tree_check.cpp"),
                    os.path.join(ROOT, "box2d_amd", "csrc", "tree.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("n", [50, 2000, 12000])
def test_deferred_enlargement_builds_the_same_tree(tree_check, n):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    proc = subprocess.run([tree_check, str(n)], capture_output=True, text=True, env=env)
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert proc.stderr == "", proc.stderr[-2000:]
    out = proc.stdout.split()
    assert len(out) == 3
    assert out[0] == out[1] == out[2], f"immediate {out[0]}, deferred {out[1]}, deferred reversed {out[2]}"
