"""Pins the oracle's Ball-vs-convex contact (contact_ball_convex_polyhedron.rs:12-63)
and the point projections under it, against computations that share no code
with it:

  * Cuboid (point_aabb.rs): clamping the ball centre to the box (outside) and
    the nearest face plane (inside), in f64;
  * Capsule (point_capsule.rs): distance to the axis segment minus the radius;
  * Cone (point_cone.rs): the 2D distance to the cone's half cross-section in
    the plane through the axis and the centre;
  * ConvexPolyhedron (point_support_map.rs, GJK + EPA on ConstantOrigin):
    brute force over the hull's facets (point-triangle distances outside,
    facet-plane distances inside);
  * known answers for the degenerate branch (centre on the surface), the
    Ball-first order (the flipped contact of the same pair), and f32 vs f64.
"""
import numpy as np
import pytest

from parry_amd import SHAPE_DESC_DTYPE, scenes

SH_BALL, SH_CUBOID, SH_CAPSULE, SH_CONE, SH_POLY = 0, 1, 2, 3, 4
IDENT = [0.0, 0.0, 0.0, 1.0]


def desc(kind, first=0, count=0, params=(0, 0, 0, 0)):
    d = np.zeros(1, SHAPE_DESC_DTYPE)
    d["type"], d["first_vertex"], d["num_vertices"] = kind, first, count
    d["params"] = params
    return d


def quat_matrix(q):
    x, y, z, w = q
    return np.array(
        [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _local_centre(pose_shape, pose_ball):
    R = quat_matrix(pose_shape[:4].astype(np.float64))
    return R.T @ (pose_ball[4:].astype(np.float64) - pose_shape[4:].astype(np.float64))


def _random_pairs(rng, n, spread):
    q = scenes.random_rotations(2 * n, rng).astype(np.float32)
    t = rng.uniform(-spread, spread, size=(2 * n, 3)).astype(np.float32)
    t[1::2] = 0.0  # shape at the origin, ball around it
    poses = np.concatenate([q, t], axis=1)
    return poses


def _run(orc, descs, verts, poses, prediction=100.0, f64=False):
    m = len(descs) // 2
    pairs = np.stack([np.arange(m) * 2 + 1, np.arange(m) * 2], axis=1).astype(np.uint32)  # (shape, ball)
    pairs_rev = pairs[:, ::-1].copy()  # (ball, shape)
    dt = np.float64 if f64 else np.float32
    r = orc.contact_pairs(descs, verts.astype(dt), poses.astype(dt), pairs, prediction, f64=f64)
    rr = orc.contact_pairs(descs, verts.astype(dt), poses.astype(dt), pairs_rev, prediction, f64=f64)
    return r, rr


def _check_flip(r, rr, tol, radii):
    """(shape, ball) and (ball, shape) give the same contact with the roles swapped.

    The normal is (projection - centre) / |projection - centre|, so a rounding
    difference in the centre moves it by about tol / |projection - centre|: the
    normals (and the ball's contact point, which follows them) are compared
    with that allowance."""
    assert (r["status"] == 1).all() and (rr["status"] == 1).all()
    o, q = r["out"].astype(np.float64), rr["out"].astype(np.float64)
    assert np.allclose(o[:, 12], q[:, 12], rtol=0, atol=tol)
    assert np.allclose(o[:, 0:3], q[:, 3:6], atol=tol)
    gap = np.abs(o[:, 12] + radii.astype(np.float64))  # |projection - centre|
    allow = tol + 4 * tol / np.maximum(gap, 1e-6)
    for a, b in ((o[:, 3:6], q[:, 0:3]), (o[:, 6:9], q[:, 9:12]), (o[:, 9:12], q[:, 6:9])):
        assert (np.abs(a - b).max(axis=1) <= allow * np.maximum(1.0, radii)).all()


def _interleave(shape_descs, radii):
    out = []
    for d, r in zip(shape_descs, radii):
        out.append(d)
        out.append(desc(SH_BALL, params=(r, 0, 0, 0)))
    return np.concatenate(out)


def test_ball_cuboid_closed_form(orc):
    rng = np.random.default_rng(11)
    n = 400
    he = rng.uniform(0.3, 1.5, size=(n, 3)).astype(np.float32)
    radii = rng.uniform(0.1, 0.8, size=n).astype(np.float32)
    descs = []
    for i in range(n):
        descs.append(desc(SH_BALL, params=(radii[i], 0, 0, 0)))
        descs.append(desc(SH_CUBOID, params=(*he[i], 0)))
    descs = np.concatenate(descs)
    poses = _random_pairs(rng, n, 2.5)
    r, rr = _run(orc, descs, np.zeros((0, 3), np.float32), poses)
    inside = 0
    for i in range(n):
        c = _local_centre(poses[2 * i + 1], poses[2 * i])
        h = he[i].astype(np.float64)
        q = np.clip(c, -h, h)
        if np.any(q != c):
            expect = np.linalg.norm(c - q) - radii[i]
        else:
            inside += 1
            expect = -np.min(h - np.abs(c)) - radii[i]
        assert abs(r["out"][i, 12] - expect) <= 2e-5 * max(1.0, abs(expect)), (i, r["out"][i, 12], expect)
        # point1 lies on the cuboid surface (local frame)
        R = quat_matrix(poses[2 * i + 1, :4].astype(np.float64))
        p1 = R.T @ (r["out"][i, 0:3].astype(np.float64) - poses[2 * i + 1, 4:])
        assert np.max(np.abs(p1) - h) <= 2e-5 and np.min(h - np.abs(p1)) <= 2e-5
    assert inside > 10
    _check_flip(r, rr, 2e-5, radii)


def test_ball_capsule_closed_form(orc):
    rng = np.random.default_rng(12)
    n = 400
    hh = rng.uniform(0.2, 1.2, size=n).astype(np.float32)
    rc = rng.uniform(0.2, 0.7, size=n).astype(np.float32)
    radii = rng.uniform(0.1, 0.8, size=n).astype(np.float32)
    descs = []
    for i in range(n):
        descs.append(desc(SH_BALL, params=(radii[i], 0, 0, 0)))
        descs.append(desc(SH_CAPSULE, params=(hh[i], rc[i], 0, 0)))
    descs = np.concatenate(descs)
    poses = _random_pairs(rng, n, 2.0)
    r, rr = _run(orc, descs, np.zeros((0, 3), np.float32), poses)
    inside = 0
    for i in range(n):
        c = _local_centre(poses[2 * i + 1], poses[2 * i])
        q = np.array([0.0, np.clip(c[1], -hh[i], hh[i]), 0.0])
        d = np.linalg.norm(c - q)
        inside += d < rc[i]
        expect = d - rc[i] - radii[i]
        assert abs(r["out"][i, 12] - expect) <= 2e-5 * max(1.0, abs(expect)), (i, r["out"][i, 12], expect)
    assert inside > 10
    _check_flip(r, rr, 2e-5, radii)


def _seg_dist2d(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def cone_signed_distance(c, h, rad):
    """Signed distance from c to a y-axis cone (apex (0,h,0), base disc at y=-h), f64."""
    rho, y = np.hypot(c[0], c[2]), c[1]
    p = np.array([rho, y])
    apex, rim, base = np.array([0.0, h]), np.array([rad, -h]), np.array([0.0, -h])
    d_side, d_base = _seg_dist2d(p, apex, rim), _seg_dist2d(p, rim, base)
    # inside test: below the side line and above the base
    side_n = np.array([2 * h, rad])  # outward normal of the slanted side in (rho, y)
    inside = y >= -h and np.dot(p - apex, side_n) <= 0.0
    d = min(d_side, d_base)
    return -d if inside else d


def test_ball_cone_cross_section(orc):
    rng = np.random.default_rng(13)
    n = 400
    hh = rng.uniform(0.3, 1.2, size=n).astype(np.float32)
    rad = rng.uniform(0.3, 1.0, size=n).astype(np.float32)
    radii = rng.uniform(0.1, 0.6, size=n).astype(np.float32)
    descs = []
    for i in range(n):
        descs.append(desc(SH_BALL, params=(radii[i], 0, 0, 0)))
        descs.append(desc(SH_CONE, params=(hh[i], rad[i], 0, 0)))
    descs = np.concatenate(descs)
    poses = _random_pairs(rng, n, 1.0)
    r, rr = _run(orc, descs, np.zeros((0, 3), np.float32), poses)
    inside = 0
    for i in range(n):
        c = _local_centre(poses[2 * i + 1], poses[2 * i])
        sd = cone_signed_distance(c, float(hh[i]), float(rad[i]))
        inside += sd < 0
        expect = sd - radii[i]
        assert abs(r["out"][i, 12] - expect) <= 3e-5 * max(1.0, abs(expect)), (i, c, r["out"][i, 12], expect)
    assert inside > 10
    _check_flip(r, rr, 3e-5, radii)


def _point_triangle_distance(p, a, b, c):
    """Closest-point distance by minimising over the triangle (barycentric clamp), f64."""
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric coordinates of q
    v0, v1, v2 = ab, ac, q - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    if v >= 0 and w >= 0 and v + w <= 1:
        return abs(np.dot(p - a, n))
    return min(_seg3(p, a, b), _seg3(p, b, c), _seg3(p, c, a))


def _seg3(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def test_ball_polyhedron_brute_force(orc):
    from scipy.spatial import ConvexHull

    rng = np.random.default_rng(14)
    n = 200
    verts, descs, hulls, radii = [], [], [], rng.uniform(0.1, 0.6, size=n).astype(np.float32)
    first = 0
    for i in range(n):
        pts = scenes.convex_hull_vertices(scenes.convex_point_cloud(rng, int(rng.integers(8, 65))))
        pts = np.asarray(pts, np.float32)
        hulls.append(ConvexHull(pts.astype(np.float64)))
        descs.append(desc(SH_BALL, params=(radii[i], 0, 0, 0)))
        descs.append(desc(SH_POLY, first, pts.shape[0]))
        verts.append(pts)
        first += pts.shape[0]
    descs = np.concatenate(descs)
    verts = np.concatenate(verts)
    poses = _random_pairs(rng, n, 1.2)
    r, rr = _run(orc, descs, verts, poses)
    inside = 0
    for i in range(n):
        c = _local_centre(poses[2 * i + 1], poses[2 * i])
        hull = hulls[i]
        planes = hull.equations  # n . x + d <= 0 inside
        sd_planes = planes[:, :3] @ c + planes[:, 3]
        if np.all(sd_planes <= 0):
            inside += 1
            expect = np.max(sd_planes) - radii[i]
        else:
            P = hull.points
            expect = min(_point_triangle_distance(c, *P[s]) for s in hull.simplices) - radii[i]
        # GJK stops once its bounds agree to eps_rel = sqrt(10 EPS) ~ 1.1e-3 of
        # the distance (gjk.rs:367, 419-424): parry's own accuracy outside.
        tol = 5e-5 if np.all(sd_planes <= 0) else 1.1e-3 * (expect + radii[i]) + 5e-5
        assert abs(r["out"][i, 12] - expect) <= tol, (i, r["out"][i, 12], expect)
    assert inside > 10
    _check_flip(r, rr, 5e-5, radii)


def test_polyhedron_box_corners_match_cuboid(orc):
    rng = np.random.default_rng(15)
    n = 200
    he = rng.uniform(0.3, 1.2, size=(n, 3)).astype(np.float32)
    radii = rng.uniform(0.1, 0.6, size=n).astype(np.float32)
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    verts, dpoly, dbox = [], [], []
    for i in range(n):
        verts.append(corners * he[i])
        dpoly += [desc(SH_BALL, params=(radii[i], 0, 0, 0)), desc(SH_POLY, 8 * i, 8)]
        dbox += [desc(SH_BALL, params=(radii[i], 0, 0, 0)), desc(SH_CUBOID, params=(*he[i], 0))]
    poses = _random_pairs(rng, n, 2.0)
    rp, _ = _run(orc, np.concatenate(dpoly), np.concatenate(verts), poses)
    rb, _ = _run(orc, np.concatenate(dbox), np.zeros((0, 3), np.float32), poses)
    assert np.allclose(rp["out"][:, 12], rb["out"][:, 12], atol=3e-5)
    assert np.allclose(rp["out"][:, 6:9], rb["out"][:, 6:9], atol=1e-3)


def test_degenerate_centre_on_surface_kats(orc):
    """Ball centre exactly on the surface: dist = -radius and the feature normal
    (Cuboid) or the fallback normal (point_support_map.rs / shape.rs defaults)."""
    r = 0.25
    cases = [  # (shape desc, ball centre in the shape frame, expected normal1)
        (desc(SH_CUBOID, params=(1.0, 0.5, 0.5, 0)), (1.0, 0.25, -0.125), (1.0, 0.0, 0.0)),
        (desc(SH_CUBOID, params=(1.0, 0.5, 0.5, 0)), (0.25, -0.5, 0.125), (0.0, -1.0, 0.0)),
        # a centre on a corner has a zero shift on every axis, so point_aabb.rs:83-92
        # reports the first face it lies on (Face(3) = -x), not the vertex
        (desc(SH_CUBOID, params=(1.0, 0.5, 0.5, 0)), (-1.0, 0.5, 0.5), (-1.0, 0.0, 0.0)),
    ]
    for d, c, n1 in cases:
        descs = np.concatenate([d, desc(SH_BALL, params=(r, 0, 0, 0))])
        poses = np.array([IDENT + [0, 0, 0], IDENT + list(c)], np.float32)
        res = orc.contact_pairs(descs, np.zeros((0, 3), np.float32), poses, [[0, 1]], 0.01)
        assert res["status"][0] == 1
        assert res["out"][0, 12] == np.float32(-r)
        assert np.allclose(res["out"][0, 6:9], n1, atol=1e-7), (c, res["out"][0, 6:9])
        assert tuple(res["out"][0, 0:3]) == tuple(np.float32(c))


def test_capsule_axis_centre_kat(orc):
    """Centre on the capsule's axis: the projection uses the axis' orthonormal
    basis vector (+x for a y-axis segment), so normal1 = +x, dist = -(rc + r)."""
    descs = np.concatenate([desc(SH_CAPSULE, params=(1.0, 0.5, 0, 0)), desc(SH_BALL, params=(0.25, 0, 0, 0))])
    poses = np.array([IDENT + [0, 0, 0], IDENT + [0, 0.25, 0]], np.float32)
    res = orc.contact_pairs(descs, np.zeros((0, 3), np.float32), poses, [[0, 1]], 0.01)
    assert res["status"][0] == 1
    assert res["out"][0, 12] == np.float32(-0.75)
    assert tuple(res["out"][0, 6:9]) == (1.0, 0.0, 0.0)
    assert tuple(res["out"][0, 0:3]) == (0.5, 0.25, 0.0)


def test_mixed_f32_matches_f64(orc):
    descs, verts = scenes.mixed_shapes(4096, seed=21)
    rng = np.random.default_rng(5)
    balls = np.nonzero(descs["type"] == SH_BALL)[0]
    others = np.nonzero(descs["type"] != SH_BALL)[0]
    m = min(len(balls), len(others))
    i, j = rng.permutation(balls)[:m], rng.permutation(others)[:m]
    pairs = np.stack([i, j], axis=1).astype(np.uint32)
    pairs[::2] = pairs[::2, ::-1]
    q = scenes.random_rotations(len(descs), rng).astype(np.float32)
    t = np.zeros((len(descs), 3), np.float32)
    t[j] = rng.uniform(-2.5, 2.5, size=(m, 3))
    poses = np.concatenate([q, t], axis=1)
    r32 = orc.contact_pairs(descs, verts, poses, pairs, 0.01)
    r64 = orc.contact_pairs(descs, verts.astype(np.float64), poses.astype(np.float64), pairs, 0.01, f64=True)
    assert (r32["status"] >= 0).all()
    both = (r32["status"] == 1) & (r64["status"] == 1)
    assert both.sum() > 100
    assert (r32["status"] != r64["status"]).sum() <= 0.002 * len(pairs)
    err = np.abs(r32["out"][both, 12] - r64["out"][both, 12])
    print(f"ball-convex f32 vs f64: {both.sum()} contacts, max |dist err| {err.max():.3e}")
    # closed forms agree to f32 rounding; polyhedra run GJK, whose stopping
    # tolerance sqrt(10 EPS) differs between f32 (1.1e-3) and f64 (4.7e-8)
    assert np.quantile(err, 0.9) <= 1e-5 and err.max() <= 2e-3
