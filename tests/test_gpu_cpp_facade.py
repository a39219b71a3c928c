"""The C++ RayTracer-shaped facade (include/xdg_hip.hpp), run as its own
program on the GPU: MeshMock known answers from the reference's tests,
and the mesh-manager flattening header run against test managers."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_ray_tracer")


def test_cpp_facade_mesh_mock():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_ray_tracer is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout


FLATTEN = os.path.join(ROOT, "tests", "cpp", "test_mesh_flatten")


def test_mesh_manager_flattening():
    """include/xdg_hip_mesh.hpp over mesh managers shaped like
    xdg::MeshManager (tests/cpp/test_mesh_flatten.cpp): id remaps, the
    reverse-sense flip, boundary faces and re-entry, a global element tree,
    MOAB versus mock face order."""
    if not os.path.exists(FLATTEN):
        pytest.fail("tests/cpp/test_mesh_flatten is not built (run __graft_entry__.build())")
    r = subprocess.run([FLATTEN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
