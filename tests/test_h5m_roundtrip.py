""".h5m geometry files (xdg_amd/h5m.py over libxdg_h5m.so), CPU only.

The synthetic configurations are written as DAGMC-style MOAB files and read
back: coordinates, each surface's triangles, the GEOM_SENSE_2 parents, the
volumes, each volume's tetrahedra and the groups come back unchanged, the
implicit complement is rebuilt as MeshManager::create_implicit_complement
does, and queries on trees built from the read model give the same bits as
on the original.  No MOAB-written .h5m exists here, so compatibility with
files written by MOAB itself is unpinned: the layout is MOAB's /tstt layout
as documented in include/xdg_h5m.h.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from xdg_amd import geometry as G
from xdg_amd import h5m
from xdg_amd import philox as P

pytestmark = pytest.mark.skipif(not os.path.exists(h5m.LIB_PATH), reason="libxdg_h5m.so not built (no HDF5)")


def _same_surfaces(a, b):
    assert sorted(a.surfaces) == sorted(b.surfaces)
    for s in a.surfaces:
        ta = a.verts[a.tris[a.surface_faces(s)]]
        tb = b.verts[b.tris[b.surface_faces(s)]]
        assert np.array_equal(ta, tb), s


def test_single_sphere_round_trip(tmp_path):
    m = G.single_sphere(4, 10.0)
    path = tmp_path / "sphere.h5m"
    h5m.write_model(path, m)
    r = h5m.read_model(path)
    assert np.array_equal(r.model.verts, m.verts)
    assert np.array_equal(r.model.tris, m.tris)
    _same_surfaces(m, r.model)
    assert r.model.volumes == m.volumes and r.model.ipc == m.ipc
    assert r.model.senses == m.senses
    assert r.model.bump(1) == m.bump(1)
    assert len(np.unique(r.tri_ids)) == m.tris.shape[0]


def test_nested_spheres_senses_and_queries(tmp_path):
    m = G.nested_spheres(3)
    path = tmp_path / "nested.h5m"
    h5m.write_model(path, m)
    r = h5m.read_model(path).model
    _same_surfaces(m, r)
    assert r.senses == m.senses
    assert r.volumes == m.volumes and r.ipc == m.ipc
    n = 20000
    pts = P.box_points(2, P.STREAM_POS, n, [-16.0] * 3, [16.0] * 3)
    dirs = P.iso_dirs(2, P.STREAM_DIR, n)
    for v in m.volumes:
        a = O.Tree.from_model(m, v)
        b = O.Tree.from_model(r, v)
        ta, sa, _ = a.ray_fire(pts, dirs)
        tb, sb, _ = b.ray_fire(pts, dirs)
        assert np.array_equal(ta.view(np.uint64), tb.view(np.uint64)) and np.array_equal(sa, sb)
    fa = O.find_volume([O.Tree.from_model(m, v) for v in m.volumes], m.volumes, m.ipc, pts, dirs)
    fb = O.find_volume([O.Tree.from_model(r, v) for v in r.volumes], r.volumes, r.ipc, pts, dirs)
    assert np.array_equal(fa, fb)


def test_lattice_round_trip(tmp_path):
    m = G.sphere_lattice(2, 2, 0.8, 2.0)
    path = tmp_path / "lattice.h5m"
    h5m.write_model(path, m)
    r = h5m.read_model(path).model
    _same_surfaces(m, r)
    assert r.senses == m.senses and r.volumes == m.volumes and r.ipc == m.ipc


def test_tet_mesh_and_groups(tmp_path):
    mesh = G.kuhn_cube(6, 0.2, seed=4)
    bm = mesh.boundary_model()
    vol = [v for v in bm.volumes if v != bm.ipc][0]
    groups = [("boundary:vacuum", [(2, s) for s in bm.surfaces]), ("mat:water", [(3, vol)])]
    path = tmp_path / "cube.h5m"
    h5m.write_model(path, bm, vol_tets={vol: mesh.tets}, groups=groups)
    r = h5m.read_model(path)
    _same_surfaces(bm, r.model)
    assert np.array_equal(r.vol_tets[vol], mesh.tets)
    assert np.array_equal(r.model.verts, bm.verts)
    assert r.groups == groups
    assert r.boundary_conditions() == {s: 1 for s in bm.surfaces}
    # the element walk on the read tets is the walk on the original
    rebuilt = G.TetMesh(r.model.verts, r.vol_tets[vol], mesh.adj, mesh.face_order, mesh.bfaces, mesh.bface_elem)
    p0 = P.box_points(4, P.STREAM_POS, 2000, [0.0] * 3, [1.0] * 3)
    p1 = P.box_points(4, P.STREAM_DIR, 2000, [0.0] * 3, [1.0] * 3)
    ca, ha = O.TetMesh(mesh).segments(p0, p1)
    cb, hb = O.TetMesh(rebuilt).segments(p0, p1)
    assert np.array_equal(ca, cb) and np.array_equal(ha, hb)


def test_errors(tmp_path):
    bad = tmp_path / "not_hdf5.h5m"
    bad.write_bytes(b"plain text, not HDF5")
    with pytest.raises(h5m.H5mError, match="error -1"):
        h5m.read_model(bad)
    with pytest.raises(h5m.H5mError):
        h5m.read_model(tmp_path / "missing.h5m")
    m = G.single_sphere(1, 1.0)
    m2 = G.SurfaceModel(m.verts, m.tris, m.tri_surf, m.surfaces, {1: (7, 2)}, [1, 2], ipc=2)
    with pytest.raises(h5m.H5mError, match="unknown volume"):
        h5m.write_model(tmp_path / "bad_sense.h5m", m2)


def test_h5m_mesh_manager_flattening(tmp_path):
    """A .h5m file through the C++ H5mMeshManager (include/xdg_h5m.hpp) and
    the flattening header: the element arrays, the adjacency found by face
    matching, the boundary-face map and the flip flags equal the mesh's own
    (tests/cpp/h5m_flatten_dump.cpp writes them)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "dump"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(root, "include"), os.path.join(root, "tests", "cpp", "h5m_flatten_dump.cpp"),
                    "-L", os.path.dirname(h5m.LIB_PATH), "-lxdg_h5m",
                    "-Wl,-rpath," + os.path.dirname(h5m.LIB_PATH), "-o", str(exe)], check=True)
    mesh = G.kuhn_cube(5, 0.2, seed=4)
    bm = mesh.boundary_model()
    vol = [v for v in bm.volumes if v != bm.ipc][0]
    path = tmp_path / "cube.h5m"
    h5m.write_model(path, bm, vol_tets={vol: mesh.tets})
    out = tmp_path / "out"
    out.mkdir()
    subprocess.run([str(exe), str(path), str(out)], check=True)
    ld = lambda name, dt: np.fromfile(out / name, dtype=dt)  # noqa: E731
    ids = ld("vol_ids.i32", np.int32)
    assert ids.tolist() == [vol, bm.ipc]
    tets = ld("tets.i32", np.int32).reshape(-1, 4)
    verts = ld("verts.f64", np.float64).reshape(-1, 3)
    # the manager's local vertex numbering: the same tets, vertex by vertex
    assert np.array_equal(verts[tets], mesh.verts[mesh.tets])
    assert np.array_equal(ld("adj.i32", np.int32).reshape(-1, 4), mesh.adj)
    elem_ids = ld("elem_ids.i32", np.int32)
    assert len(np.unique(elem_ids)) == mesh.tets.shape[0]
    # boundary faces: file ids of the complement's faces -> the owning tet
    r = h5m.read_model(path)
    faces = ld("ipc_faces.i32", np.int32)
    bmap = ld("bmap.i32", np.int32)
    tri_of = {int(f): k for k, f in enumerate(r.tri_ids)}
    owner = {tuple(sorted(t)): e for e, t in zip(mesh.bface_elem, mesh.bfaces.tolist())}
    for f in faces:
        tri = r.model.tris[tri_of[int(f)]]
        assert bmap[f] == owner[tuple(sorted(tri.tolist()))]
    assert (ld("vol_flip.u8", np.uint8) == 0).all() and (ld("ipc_flip.u8", np.uint8) == 1).all()
    assert np.array_equal(ld("bumps.f64", np.float64), [bm.bump(vol), bm.bump(bm.ipc)])
