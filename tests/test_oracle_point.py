"""Oracle pins for the TriMesh point projection and the culling ray cast.

Point projection (point_composite_shape.rs:59-80, point_triangle.rs:54-291):
  * closed forms on one triangle: the foot of the perpendicular over the face,
    the vertex and edge Voronoi regions, a point on an edge projecting to itself
    (to f32 rounding) for every rotation of the vertices (the 3D form of
    parry2d's project_local_point_point_on_ab test), is_inside as relative_eq!, and the
    solid / closest-edge fallbacks of a degenerate (collinear) triangle;
  * the Bvh traversal against brute force over all triangles: the same smallest
    distance (the find_best cost) for every point.
Culling cast (ray_trimesh.rs:58-65, 131-178):
  * the reference's own cast_ray_on_trimesh_with_culling test (ray_trimesh.rs:187-212);
  * the Bvh traversal against brute force over the triangles the culling keeps;
  * rays from outside a closed outward-wound mesh: ignoring back faces changes
    nothing (every first hit is a front face), ignoring front faces finds the
    exit face.
"""
import numpy as np

from oracle import oracle as orc
from parry_amd import scenes


def _f32_dist(p, q):
    d = (np.asarray(p, np.float32) - np.asarray(q, np.float32)).astype(np.float32)
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return np.sqrt(s.astype(np.float32))


TRI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)


def test_triangle_projection_closed_forms():
    p, ins = orc.triangle_project_point(TRI, [0.25, 0.5, 3.0], False)
    assert tuple(p) == (0.25, 0.5, 0.0) and not ins  # foot of the perpendicular
    p, ins = orc.triangle_project_point(TRI, [0.25, 0.5, 0.0], False)
    assert tuple(p) == (0.25, 0.5, 0.0) and ins  # on the face
    assert tuple(orc.triangle_project_point(TRI, [-1.0, -2.0, 0.5], False)[0]) == (0, 0, 0)  # vertex a
    assert tuple(orc.triangle_project_point(TRI, [3.0, -0.5, 1.0], False)[0]) == (1, 0, 0)  # vertex b
    assert tuple(orc.triangle_project_point(TRI, [-0.5, 2.0, 0.0], False)[0]) == (0, 1, 0)  # vertex c
    assert tuple(orc.triangle_project_point(TRI, [0.5, -1.0, 2.0], False)[0]) == (0.5, 0.0, 0.0)  # edge ab
    assert tuple(orc.triangle_project_point(TRI, [-1.0, 0.25, 0.0], False)[0]) == (0.0, 0.25, 0.0)  # edge ac
    p, _ = orc.triangle_project_point(TRI, [1.0, 1.0, 0.0], False)  # edge bc, midpoint
    assert np.allclose(p, [0.5, 0.5, 0.0], atol=1e-7)
    # relative_eq!: a point 1 ulp-ish off the face still counts as inside
    _, ins = orc.triangle_project_point(TRI, [0.25, 0.5, 1e-8], False)
    assert ins


def test_point_on_edge_projects_to_itself_for_every_rotation():
    # In 3D the edge is found through a different vertex pair for each
    # rotation, so the projection is the query point up to f32 rounding
    # (1.4 comes back as 1.4000001 through edge bc), and relative_eq! calls it inside.
    v = np.array([[2.0, 1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0]], np.float32)
    q = np.array([1.4, 1.0, 0.0], np.float32)
    for k in range(3):
        p, ins = orc.triangle_project_point(np.roll(v, k, axis=0), q, False)
        assert np.all(np.abs(p - q) <= 2 * np.spacing(np.float32(1.4))) and ins
    # and off the plane: the projection is the point's foot on the edge
    p, ins = orc.triangle_project_point(v, [1.4, 1.0, 0.5], False)
    assert np.array_equal(p, q) and not ins


def test_degenerate_triangle_fallbacks():
    line = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)  # collinear: va + vb + vc == 0
    q = [0.5, 1.0, 0.0]
    p, ins = orc.triangle_project_point(line, q, True)
    assert np.array_equal(p, np.float32(q)) and ins  # solid: the point itself, inside
    p, ins = orc.triangle_project_point(line, q, False)
    assert ins and abs(p[1]) < 1e-6 and abs(p[2]) < 1e-6 and 0.0 <= p[0] <= 2.0  # on the segment


def test_bvh_projection_equals_brute_force():
    v, f = scenes.noisy_icosphere(3, seed=1)
    m = orc.OracleTriMesh(v, f)
    rng = np.random.default_rng(4)
    pts = np.concatenate([rng.uniform(-1.6, 1.6, (6000, 3)), rng.normal(0, 0.3, (2000, 3))]).astype(np.float32)
    for solid in (False, True):
        a = m.project_points(pts, solid)
        b = m.project_points(pts, solid, brute=True)
        da, db = _f32_dist(pts, a["point"]), _f32_dist(pts, b["point"])
        assert np.array_equal(da, db), f"solid={solid}: {np.count_nonzero(da != db)} distances differ"
        assert a["W"] > 0 and a["K"] < b["point"].shape[0] * f.shape[0] / 20  # the tree prunes
    # Triangle ids are not compared: a point whose projection is a vertex or
    # edge shared by several triangles is a tie, which find_best and the brute
    # force break by their own visiting orders.


def test_projection_inside_flags_on_the_surface():
    v, f = scenes.noisy_icosphere(2, seed=1)
    m = orc.OracleTriMesh(v, f)
    # triangle centroids lie on the surface: projected onto themselves, is_inside
    c = v[f].mean(axis=1).astype(np.float32)
    r = m.project_points(c, False)
    assert r["is_inside"].all()
    assert np.abs(r["point"] - c).max() < 1e-6
    far = (c * 3.0).astype(np.float32)
    assert not m.project_points(far, False)["is_inside"].any()


def test_culling_reference_kat():
    mesh = orc.OracleTriMesh(TRI, np.array([[0, 1, 2]], np.uint32))
    up = np.array([[0, 0, -1, 0, 0, 1]], np.float32)
    down = np.array([[0, 0, 1, 0, 0, -1]], np.float32)
    miss = np.uint32(0xFFFFFFFF)
    # IgnoreFrontfaces (2): up hits (a back face), down misses; IgnoreBackfaces (1): the reverse
    assert mesh.cast_rays_with_culling(up, 1000.0, 2)["feature"][0] != miss
    assert mesh.cast_rays_with_culling(down, 1000.0, 2)["feature"][0] == miss
    assert mesh.cast_rays_with_culling(up, 1000.0, 1)["feature"][0] == miss
    assert mesh.cast_rays_with_culling(down, 1000.0, 1)["feature"][0] != miss


def test_culling_bvh_equals_brute_force_and_closed_mesh_properties():
    v, f = scenes.noisy_icosphere(3, seed=1)
    m = orc.OracleTriMesh(v, f)
    rays = scenes.rays_into_unit_ball(20000, seed=5)
    plain = m.cast_rays(rays, 10.0)
    nt = f.shape[0]
    for cull in (1, 2):
        a = m.cast_rays_with_culling(rays, 10.0, cull)
        b = m.cast_rays_with_culling(rays, 10.0, cull, brute=True)
        hit = a["feature"] != 0xFFFFFFFF
        assert np.array_equal(hit, b["feature"] != 0xFFFFFFFF)
        assert np.array_equal(a["toi"][hit], b["toi"][hit])
        if cull == 1:  # from outside, first hits are front faces: nothing changes
            assert np.array_equal(a["feature"], plain["feature"]) and np.array_equal(a["toi"], plain["toi"])
            assert (a["feature"][hit] < nt).all()
        else:  # only back faces remain: the ray's exit, farther than the entry
            assert (a["feature"][hit] >= nt).all()
            both = hit & (plain["feature"] != 0xFFFFFFFF)
            assert (a["toi"][both] > plain["toi"][both]).all()
