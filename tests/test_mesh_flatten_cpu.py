"""The mesh-manager flattening header (include/xdg_hip_mesh.hpp) on the CPU:
tests/cpp/flatten_cpu.cpp checks the arrays flattened from the test managers
(flip flags, id remaps, adjacency within and across volumes, the boundary
face map, the bump).  The GPU half is tests/cpp/test_mesh_flatten.cpp."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_flatten_arrays(tmp_path):
    exe = tmp_path / "flatten_cpu"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "flatten_cpu.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
