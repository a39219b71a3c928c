"""A .h5m file to device trees on the GPU (tests/cpp/h5m_device_walk.cpp:
H5mMeshManager -> flattening header -> element tree, implicit-complement
tree and boundary map -> xdg_hip_tally_segments): the element sequences
(counts and hashes) equal the oracle's on the original mesh, segments that
start or end outside the mesh (re-entry through the complement) included."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import h5m
from xdg_amd import philox as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "h5m_device_walk")


def test_h5m_file_to_device_walk(tmp_path):
    if not os.path.exists(h5m.LIB_PATH):
        pytest.skip("libxdg_h5m.so not built (no HDF5)")
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/h5m_device_walk is not built (run __graft_entry__.build())")
    mesh = G.kuhn_cube(10, 0.2, seed=4)
    bm = mesh.boundary_model()
    vol = [v for v in bm.volumes if v != bm.ipc][0]
    path = tmp_path / "cube.h5m"
    h5m.write_model(path, bm, vol_tets={vol: mesh.tets})
    n = 20000
    p0 = P.box_points(7, P.STREAM_POS, n, [-0.2] * 3, [1.2] * 3)
    p1 = P.box_points(7, P.STREAM_DIR, n, [-0.2] * 3, [1.2] * 3)
    seg = tmp_path / "seg.f64"
    np.ascontiguousarray(np.concatenate([p0, p1], axis=1)).tofile(seg)
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(path), str(seg), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    cnt = np.frombuffer(raw[:4 * n], np.int32)
    hsh = np.frombuffer(raw[4 * n:], np.uint64)
    rc, rh = O.TetMesh(mesh).segments(p0, p1)
    assert np.array_equal(cnt, rc)
    assert np.array_equal(hsh, rh)
    assert (rc == 0).sum() > 0 and (rc > 0).sum() > n // 2
