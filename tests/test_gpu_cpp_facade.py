"""The C++ RayTracer-shaped facade (include/xdg_hip.hpp), run as its own
program on the GPU: MeshMock known answers from the reference's tests."""
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
