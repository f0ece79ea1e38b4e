"""The host-side projector maps of qm_set_projectors under AddressSanitizer.

csrc/projector_maps.h builds the covered-point map, the padded tables, the
projection work list and the scatter slots on the host.  This test compiles
it with g++ -fsanitize=address,undefined into a small driver
(tests/asan/projector_maps_check.cpp) that checks every map against its
definition on random projector sets (wrapped spheres, overlaps, atoms without
voxels, 32 projectors, points covered by more atoms than a scatter block
stages) and runs it on the CPU.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "qumasun-workshop-template_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_projector_maps_under_asan(tmp_path):
    exe = str(tmp_path / "projector_maps_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           f"-I{CSRC}", os.path.join(HERE, "asan", "projector_maps_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"build failed:\n{' '.join(cmd)}\n{r.stdout}{r.stderr}"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, f"{r.stdout}{r.stderr}"
    assert "cases ok" in r.stdout
