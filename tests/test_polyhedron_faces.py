"""The face tables scenes.polyhedron_faces builds for ConvexPolyhedra (the layout of
ConvexPolyhedron's faces, vertices_adj_to_face and edges_adj_to_face,
convex_polyhedron.rs:172-185): per polyhedron, Euler's V - E + F = 2, every
edge bordering exactly two faces, in opposite directions; unit outward
normals with every vertex on or behind every face plane; each face's vertices
on its plane and counter-clockwise seen from outside; edge k of a face joining
its vertices k and k + 1; coplanar hull triangles merged (a box has six
quadrilaterals)."""
import numpy as np

from parry_amd import SHAPE_DESC_DTYPE, scenes


def _check(verts, faces, fv, fe, first, count):
    first, count = int(first), int(count)
    lo = int(faces["first"][first])
    hi = int(faces["first"][first + count - 1]) + int(faces["count"][first + count - 1])
    V = len(np.unique(fv[lo:hi]))
    edges = {}
    for f in range(first, first + count):
        a, c = int(faces["first"][f]), int(faces["count"][f])
        ids = fv[a:a + c]
        n = faces["normal"][f].astype(np.float64)
        assert abs(np.linalg.norm(n) - 1.0) < 1e-6
        pts = verts[ids].astype(np.float64)
        off = float(np.dot(n, pts[0]))
        assert np.abs(pts @ n - off).max() < 1e-5  # on its plane
        assert (verts.astype(np.float64) @ n - off).max() < 1e-5  # every vertex behind it
        c0 = pts.mean(0)
        for k in range(c):  # counter-clockwise around the outward normal
            assert np.dot(np.cross(pts[k] - c0, pts[(k + 1) % c] - c0), n) > 0
        for k in range(c):
            e = (int(ids[k]), int(ids[(k + 1) % c]))
            key = (min(e), max(e))
            edges.setdefault(key, []).append((int(fe[a + k]), e))
    for key, uses in edges.items():
        assert len(uses) == 2 and uses[0][0] == uses[1][0], f"edge {key}: {uses}"
        assert uses[0][1] == uses[1][1][::-1]  # opposite directions in the two faces
    E = len(edges)
    assert sorted(u[0][0] for u in edges.values()) == list(range(E))  # ids 0..E-1
    assert V - E + count == 2


def test_random_hulls_and_boxes():
    verts, offs = scenes.convex_polyhedra(200, seed=4, kmin=4, kmax=40)
    d = np.zeros(len(offs) - 1, SHAPE_DESC_DTYPE)
    d["type"] = 4
    d["first_vertex"] = offs[:-1]
    d["num_vertices"] = offs[1:] - offs[:-1]
    t = scenes.polyhedron_faces(d, verts)
    for i in range(len(d)):
        v = verts[offs[i]:offs[i + 1]]
        _check(v, t["faces"], t["face_vertices"], t["face_edges"], *t["shape_faces"][i])
    box = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32) * [0.5, 1, 2]
    db = np.zeros(2, SHAPE_DESC_DTYPE)
    db["type"] = [1, 4]  # a Cuboid gets no faces
    db["first_vertex"][1] = 0
    db["num_vertices"][1] = 8
    t = scenes.polyhedron_faces(db, box)
    assert tuple(t["shape_faces"][0]) == (0, 0) and tuple(t["shape_faces"][1]) == (0, 6)
    assert (t["faces"]["count"] == 4).all()
    _check(box, t["faces"], t["face_vertices"], t["face_edges"], 0, 6)
