"""The host query planner (fsig_amd/csrc/plan.cc) on random queries, built
with g++ under AddressSanitizer and UndefinedBehaviorSanitizer (host code
only): equitable classes with exact thresholds and neighbour-label
requirements, signature classes for alg 2, join orders that are
permutations, symmetry-breaking conditions against brute-force
automorphism groups (tests/plan_check.cc)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(HIP_INC), reason="g++ or HIP headers absent")
def test_planner_invariants_on_random_queries(tmp_path):
    exe = str(tmp_path / "plan_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-D__HIP_PLATFORM_AMD__", "-I", HIP_INC, "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "fsig_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "plan_check.cc"), os.path.join(ROOT, "fsig_amd", "csrc", "plan.cc")],
                   check=True, timeout=300)
    for seed in (1, 2):
        r = subprocess.run([exe, "20000", str(seed)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "checked 20000 queries" in r.stdout
