"""The tiled DPD permutation's host plan on the CPU (no GPU): tests/sort_plan_check.cpp
replays the kernels' index arithmetic over every work item of plan_sort's tasks and
checks bounds, LDS use, one write per output element and the source element
against fetch_addr, for 1/2/4/8 irreps, every symmetry and all 24 permutations.
The GPU tests check the kernels themselves (every sort bit-equal to the oracle)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "revised-cluster-perturbation-single-double_amd", "csrc")
HIP_INC = "/opt/rocm/include"


def test_sort_plan_covers_every_element_once(tmp_path):
    cxx = shutil.which("g++")
    if not cxx or not os.path.isdir(HIP_INC):
        pytest.skip("g++ or the HIP headers are not available")
    exe = tmp_path / "sort_plan_check"
    r = subprocess.run([cxx, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{CSRC}",
                        f"-I{os.path.join(ROOT, 'include')}", f"-I{HIP_INC}",
                        os.path.join(ROOT, "tests", "sort_plan_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    cases, _, failures, _ = r.stdout.split()[-4:]
    assert int(cases) > 2000 and int(failures) == 0, r.stdout
