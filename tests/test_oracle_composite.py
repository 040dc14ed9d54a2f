"""Oracle pins for contact between a TriMesh and convex shapes
(contact_composite_shape_shape.rs:14-59 with DefaultQueryDispatcher::contact).

  * closed forms on a flat two-triangle plane: a Ball, a Cuboid, a Capsule and a
    Cone resting across it (dist = lowest point's height, normal1 = +z);
  * Ball against a Triangle (contact_convex_polyhedron_ball with the triangle's
    closed-form projection) equals the float64 point-to-triangle distance minus
    the radius, in both orders (Triangle first, Ball first flipped);
  * the Bvh candidate selection: the same smallest dist as the contact with
    every triangle of the mesh taken one by one (brute force), for mixed shapes
    around an icosphere;
  * f32 against f64 (parry3d against parry3d-f64).
"""
import numpy as np

from oracle import oracle as orc
from parry_amd import scenes
from parry_amd.shape import SHAPE_DESC_DTYPE

PLANE_V = np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32)
PLANE_F = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
IDENT = np.array([0, 0, 0, 1, 0, 0, 0], np.float32)


def _descs(types, params):
    d = np.zeros(len(types), SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, p in enumerate(params):
        d["params"][i][: len(p)] = p
    return d


def test_plane_closed_forms():
    m = orc.OracleTriMesh(PLANE_V, PLANE_F)
    # Ball r=0.5 at z=0.4; Cuboid he=0.2 at z=0.15; Capsule (h=0.3, r=0.1) and Cone (h=0.5,
    # r=0.4), whose axes are local +y, stood up along +z (90 degrees about x) at z=0.35 and
    # z=0.45: each reaches 0.1 or 0.05 below the plane
    s = np.float32(np.sqrt(0.5))
    d = _descs([0, 1, 2, 3], [[0.5], [0.2, 0.2, 0.2], [0.3, 0.1], [0.5, 0.4]])
    poses = np.array([[0, 0, 0, 1, 0.3, 0.2, 0.4], [0, 0, 0, 1, 0.1, -0.3, 0.15],
                      [s, 0, 0, s, -1.0, 1.0, 0.35], [s, 0, 0, s, 1.5, 1.5, 0.45]], np.float32)
    r = m.contact_shapes(IDENT, d, np.zeros((0, 3), np.float32), poses, 0.01)
    assert (r["status"] == 1).all()
    # EPA stops within eps_tol = 100 EPS of a curved surface (the capsule's round end)
    np.testing.assert_allclose(r["out"][:, 12], [-0.1, -0.05, -0.05, -0.05], atol=2e-5)
    np.testing.assert_allclose(r["out"][:, 6:9], np.tile([0, 0, 1], (4, 1)), atol=1e-5)
    np.testing.assert_allclose(r["out"][:, 9:12], np.tile([0, 0, -1], (4, 1)), atol=1e-5)
    # the contact point on the mesh lies on the plane
    assert np.abs(r["out"][:, 2]).max() < 1e-5
    # a lifted mesh pose moves everything by the same offset
    lifted = np.array([0, 0, 0, 1, 0, 0, 0.25], np.float32)
    r2 = m.contact_shapes(lifted, d, np.zeros((0, 3), np.float32), poses + [0, 0, 0, 0, 0, 0, 0.25], 0.01)
    np.testing.assert_allclose(r2["out"][:, 12], r["out"][:, 12], atol=1e-6)
    # beyond prediction: None
    far = poses + [0, 0, 0, 0, 0, 0, 1.0]
    assert (m.contact_shapes(IDENT, d, np.zeros((0, 3), np.float32), far, 0.01)["status"] == 0).all()


def _point_triangle_distance64(p, a, b, c):
    # closest point on the triangle by the region tests, in float64 (independent form)
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return np.linalg.norm(p - a)
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return np.linalg.norm(p - b)
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return np.linalg.norm(p - (a + ab * (d1 / (d1 - d3))))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return np.linalg.norm(p - c)
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return np.linalg.norm(p - (a + ac * (d2 / (d2 - d6))))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        return np.linalg.norm(p - (b + (c - b) * w))
    den = 1.0 / (va + vb + vc)
    return np.linalg.norm(p - (a + ab * (vb * den) + ac * (vc * den)))


def test_ball_triangle_against_point_distance():
    rng = np.random.default_rng(2)
    for _ in range(200):
        tri = rng.normal(0, 1, (3, 3)).astype(np.float32)
        m = orc.OracleTriMesh(tri, np.array([[0, 1, 2]], np.uint32))
        c = rng.normal(0, 1, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.1, 1.0))
        pose = np.concatenate([[0, 0, 0, 1], c]).astype(np.float32)
        res = m.contact_shapes(IDENT, _descs([0], [[r]]), np.zeros((0, 3), np.float32), pose[None], 10.0)
        ref = _point_triangle_distance64(c.astype(np.float64), *tri.astype(np.float64)) - float(r)
        assert res["status"][0] == 1
        assert abs(float(res["out"][0, 12]) - ref) <= 2e-5 * max(1.0, abs(ref))


def test_bvh_selection_equals_brute_force():
    v, f = scenes.noisy_icosphere(3, seed=1)
    v = (v * 2.0).astype(np.float32)
    m = orc.OracleTriMesh(v, f)
    descs, verts, poses = scenes.shapes_around_mesh(600, 13)
    pred = 0.05
    res = m.contact_shapes(IDENT, descs, verts, poses, pred)
    # brute force: every triangle as a one-triangle mesh, the smallest dist wins
    best = np.full(len(descs), np.inf)
    for t in range(f.shape[0]):
        mt = orc.OracleTriMesh(v[f[t]], np.array([[0, 1, 2]], np.uint32))
        r = mt.contact_shapes(IDENT, descs, verts, poses, pred)
        ok = r["status"] == 1
        best[ok] = np.minimum(best[ok], r["out"][ok, 12])
    have = res["status"] == 1
    assert np.array_equal(have, np.isfinite(best))
    assert np.array_equal(res["out"][have, 12], best[have].astype(np.float32))
    assert have.sum() > 100 and (~have).sum() > 50
    assert res["candidates"] < len(descs) * f.shape[0] / 20  # the tree prunes


def test_f32_agrees_with_f64():
    v, f = scenes.noisy_icosphere(3, seed=1)
    v = (v * 2.0).astype(np.float32)
    descs, verts, poses = scenes.shapes_around_mesh(600, 17)
    r32 = orc.OracleTriMesh(v, f).contact_shapes(IDENT, descs, verts, poses, 0.05)
    r64 = orc.OracleTriMesh(v, f, f64=True).contact_shapes(IDENT, descs, verts, poses, 0.05)
    both = (r32["status"] == 1) & (r64["status"] == 1)
    assert (r32["status"] == r64["status"]).mean() > 0.99
    d = np.abs(r32["out"][both, 12] - r64["out"][both, 12])
    assert np.quantile(d, 0.99) < 1e-3
