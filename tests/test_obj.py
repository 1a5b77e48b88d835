"""Wavefront OBJ input (SimplePolygonMesh::readMeshFromObjFile,
src/SimplePolygonMesh.cpp:139-213; surface kind from SurfaceSpecifier,
src/SurfaceSpecifier.cpp:298-363).

CPU tests read files through the library's host-only entry points and check
them against a reader written here from the OBJ format; GPU tests check that
a tube or an interior boundary curve built from a file equals the one built
from the same arrays, and the mesh tube the oracle's tube of the mesh that
reader takes from the file."""
import numpy as np
import pytest

import cpmgpu
import orc
from cpmgpu import workloads


def _write(path, text):
    path.write_bytes(text.encode())
    return str(path)


def _obj_text(V, F=None, lines=None, fmt="%.17g"):
    out = ["# written by the test"]
    out += ["v " + " ".join(fmt % c for c in v) for v in V]
    if F is not None:
        out += ["f " + " ".join(str(i + 1) for i in f) for f in F]
    for ln in lines or []:
        out.append("l " + " ".join(str(i + 1) for i in ln))
    return "\n".join(out) + "\n"


def _independent_read(text):
    """Positions of "v", position indices of "f" (1-based, before any "/")."""
    V, F = [], []
    for ln in text.splitlines():
        t = ln.split()
        if not t:
            continue
        if t[0] == "v":
            V.append([float(c) for c in t[1:]])
        elif t[0] == "f":
            F.append([int(c.split("/")[0]) - 1 for c in t[1:]])
    return np.array(V), np.array(F, np.int32)


def test_triangles_with_texture_and_normal_indices(tmp_path):
    text = ("# tetrahedron\r\no tet\r\nv 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nv 0 0 1\r\nvt 0 0\r\nvt 1 0\r\n"
            "vn 0 0 1\r\ng side\r\nusemtl none\r\ns off\r\nf 1/1 3/2 2/1\r\nf 1//1 2//1 4//1\r\nf 1/1/1 4/2/1 3/1/1\r\n"
            "f 2 3 4\r\n")
    d = cpmgpu.read_obj(_write(tmp_path / "tet.obj", text))
    V, F = _independent_read(text)
    assert d["kind"] == "Triangulation"
    assert np.array_equal(d["verts"], V) and np.array_equal(d["faces"], F)


def test_icosphere_round_trip_is_exact(tmp_path):
    V, F = workloads.displaced_icosphere(3, seed=2)
    d = cpmgpu.read_obj(_write(tmp_path / "s.obj", _obj_text(V, F)))
    assert np.array_equal(d["verts"], V)            # %.17g round-trips doubles
    assert np.array_equal(d["faces"], F)


def test_closed_polyline_chain(tmp_path):
    V = workloads.torus_knot_polyline(1.0, 0.4, 50)[:-1]     # 50 distinct vertices
    n = len(V)
    d = cpmgpu.read_obj(_write(tmp_path / "k.obj", _obj_text(V, lines=[list(range(n)) + [0]])))
    assert d["kind"] == "Polyline" and d["faces"].shape == (n, 2)
    assert np.array_equal(d["faces"][:, 0], np.arange(n)) and np.array_equal(d["faces"][:, 1], (np.arange(n) + 1) % n)
    C = cpmgpu.polyline_chain(d["verts"], d["faces"])
    assert np.array_equal(C, np.vstack([V, V[:1]]))        # closed: the first vertex again


def test_polyline_pieces_and_open_curve(tmp_path):
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5]], float)
    # a curve given as several "l" records that continue each other, open
    d = cpmgpu.read_obj(_write(tmp_path / "o.obj", _obj_text(V, lines=[[0, 1], [1, 2, 3]])))
    assert d["faces"].tolist() == [[0, 1], [1, 2], [2, 3]]
    assert np.array_equal(cpmgpu.polyline_chain(d["verts"], d["faces"]), V)
    # closed through a repeated vertex (first == last vertex of the file)
    W = np.vstack([V, V[:1]])
    d = cpmgpu.read_obj(_write(tmp_path / "r.obj", _obj_text(W, lines=[[0, 1, 2, 3, 4]])))
    assert np.array_equal(cpmgpu.polyline_chain(d["verts"], d["faces"]), W)


@pytest.mark.parametrize("text,what", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3\nl 1 2\n", "only triangles or only polyline edges"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\n", "faces must be triangles"),
    ("v 0 0\nv 1 0\nl 1 2\n", "3 coordinates"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "not a positive 1-based index"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -3 -2 -1\n", "not a positive 1-based index"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "> vertex count"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n", "bad index"),
    ("# nothing\n", "no vertices"),
])
def test_rejected_files(tmp_path, text, what):
    with pytest.raises(cpmgpu.CpmError, match=what):
        cpmgpu.read_obj(_write(tmp_path / "bad.obj", text))


def test_missing_file_and_broken_chain(tmp_path):
    with pytest.raises(cpmgpu.CpmError, match="couldn't open"):
        cpmgpu.read_obj(str(tmp_path / "none.obj"))
    V = np.eye(3)
    with pytest.raises(cpmgpu.CpmError, match="consecutive"):
        cpmgpu.polyline_chain(V, np.array([[0, 1], [2, 0]], np.int32))


@pytest.mark.gpu
def test_mesh_tube_from_obj_equals_arrays(gpu_ctx, tmp_path):
    print("case obj mesh tube", flush=True)
    V, F = workloads.displaced_icosphere(3, seed=11)
    path = _write(tmp_path / "m.obj", _obj_text(V, F))
    a = cpmgpu.Tube(gpu_ctx, obj=path, dx=0.05)
    b = cpmgpu.Tube(gpu_ctx, "Triangulation", 0.05, verts=V, faces=F)
    assert a.N == b.N and a.info["xstart"] == b.info["xstart"]
    A, B = a.export(), b.export()
    for k in A:
        assert np.array_equal(A[k], B[k]), k
    # and against the oracle's tube of the mesh this test's own reader takes from the file
    Vr, Fr = _independent_read(open(path).read())
    P = orc.mesh_problem(Vr, Fr, 0.05)
    To = P.tube()
    assert a.N == P.N and a.info["xstart"] == P.xstart and a.info["N1d"] == P.N1d
    for k in ("rec", "nbr", "bdy", "bpt"):
        assert np.array_equal(A[k], To[k]), k
    assert np.abs(A["cp"] - To["cp"]).max() <= 1e-15 and np.abs(A["dist"] - To["dist"]).max() <= 1e-15
    Lg, Lo = a.L(), P.L()
    assert np.array_equal(Lg.col, Lo.col) and np.array_equal(Lg.val, Lo.val)
    Eg, Eo = a.E(), P.E()
    assert np.array_equal(Eg.col, Eo.col) and np.abs(Eg.val - Eo.val).max() <= 1e-15


@pytest.mark.gpu
def test_polyline_tube_and_ibc_from_obj(gpu_ctx, tmp_path):
    print("case obj polyline tube and ibc", flush=True)
    K = workloads.torus_knot_polyline(1.0, 0.4, 400)
    path = _write(tmp_path / "knot.obj", _obj_text(K[:-1], lines=[list(range(400)) + [0]]))
    # the curve itself as the surface (codimension two)
    a = cpmgpu.Tube(gpu_ctx, obj=path, dx=0.05)
    bb = (min(0.0, K.min()), max(0.0, K.max()))
    b = cpmgpu.Tube(gpu_ctx, "Polyline", 0.05, verts=K, bbox=bb)
    assert a.N == b.N
    A, B = a.export(), b.export()
    for k in A:
        assert np.array_equal(A[k], B[k]), k
    # the curve as an interior boundary on the torus
    tor = cpmgpu.Tube(gpu_ctx, "Torus", 0.05, params=[1.0, 0.4], bbox=(-1.4, 1.4))
    meta = dict(boundary_order=2, on_ibc_tolerance=0.1 * 0.05 ** 2)
    P1 = cpmgpu.Problem(tor, [cpmgpu.IBC.from_obj(path, **meta)])
    P2 = cpmgpu.Problem(tor, [cpmgpu.IBC("Polyline", verts=K, **meta)])
    assert P1.D == P2.D
    (L1, E1, I1), (L2, E2, I2) = P1.export(), P2.export()
    for u, v in ((L1, L2), (E1, E2)):
        assert np.array_equal(u.ptr, v.ptr) and np.array_equal(u.col, v.col) and np.array_equal(u.val, v.val)
    assert np.array_equal(I1, I2)
