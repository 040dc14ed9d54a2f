"""Independent pins for the oracle's Cuboid, Capsule and Cone functions.

Each check computes its expected value by a route that shares nothing with the
restatement under test, and each tolerance is far below the effect of a
plausible slip (a dropped radius factor, a sign flip, an unnormalised offset,
a wrong axis):

  * world AABBs (aabb_cuboid.rs:9-14 with isometry_ops.rs:14-19,
    aabb_capsule.rs:8-20, aabb_support_map.rs:12-14 -> aabb_utils.rs:10-31,
    aabb_ball.rs:25-27) against closed forms in float64 built from scipy's
    rotation matrices: |R| he for a cuboid, the posed segment +- r for a
    capsule, the apex and the rim circle's extremes c_d +- r sqrt(1 - u_d^2)
    for a cone;
  * local support points (capsule.rs:407-419, cone.rs:162-182, cuboid.rs:452-457,
    ball.rs:242-249) against the closed-form support functions
    h(d) = hh |d_y| + r |d| (capsule) and max(hh d_y, -hh d_y + r |d_xz|) (cone),
    the argmax over dense samples of the whole surface, and membership of the
    returned point in the shape's surface; plus the zero-direction and
    axial known answers;
  * Capsule-Capsule query::contact (contact_support_map_support_map.rs) against
    the segment-segment distance minus both radii, the segment distance found
    by a ternary search that uses no projection formula of the oracle;
  * Cone-Cuboid and Cone-ConvexPolyhedron distances against a constrained
    minimisation (scipy SLSQP) over the cone's own inequalities
    (|y| <= hh, |x_xz| <= r (hh - y) / (2 hh)) and the other shape's half-spaces.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from parry_amd import SHAPE_DESC_DTYPE, scenes

SH_BALL, SH_CUBOID, SH_CAPSULE, SH_CONE, SH_POLY = 0, 1, 2, 3, 4


def desc(kind, first=0, count=0, params=(0, 0, 0, 0)):
    d = np.zeros(1, SHAPE_DESC_DTYPE)
    d["type"], d["first_vertex"], d["num_vertices"] = kind, first, count
    d["params"] = params
    return d


def _rot(q):
    """Rotation matrix of a quaternion (i, j, k, w) in float64 (scipy, scalar-last)."""
    return Rotation.from_quat(np.asarray(q, np.float64)).as_matrix()


def _random_poses(rng, n, spread=3.0):
    q = scenes.random_rotations(n, rng)
    t = rng.uniform(-spread, spread, (n, 3))
    return np.concatenate([q, t], 1).astype(np.float32)


# ---- world AABBs -----------------------------------------------------------------


def test_cuboid_aabb_closed_form(orc):
    rng = np.random.default_rng(21)
    n = 400
    he = rng.uniform(0.2, 2.0, (n, 3)).astype(np.float32)
    descs = np.zeros(n, SHAPE_DESC_DTYPE)
    descs["type"] = SH_CUBOID
    descs["params"][:, :3] = he
    poses = _random_poses(rng, n)
    got = orc.shape_aabbs(descs, np.zeros((0, 3), np.float32), poses).astype(np.float64)
    for i in range(n):
        R = _rot(poses[i, :4])
        ws = np.abs(R) @ he[i].astype(np.float64)
        t = poses[i, 4:].astype(np.float64)
        np.testing.assert_allclose(got[i, :3], t - ws, atol=2e-6 * (1 + np.abs(t).max()))
        np.testing.assert_allclose(got[i, 3:], t + ws, atol=2e-6 * (1 + np.abs(t).max()))


def test_capsule_aabb_closed_form(orc):
    rng = np.random.default_rng(22)
    n = 400
    hh = rng.uniform(0.3, 1.5, n)
    r = rng.uniform(0.2, 0.8, n)
    descs = np.zeros(n, SHAPE_DESC_DTYPE)
    descs["type"] = SH_CAPSULE
    descs["params"][:, 0], descs["params"][:, 1] = hh, r
    poses = _random_poses(rng, n)
    got = orc.shape_aabbs(descs, np.zeros((0, 3), np.float32), poses).astype(np.float64)
    for i in range(n):
        R = _rot(poses[i, :4])
        t = poses[i, 4:].astype(np.float64)
        h, rad = np.float64(np.float32(hh[i])), np.float64(np.float32(r[i]))
        a, b = t + R @ [0, -h, 0], t + R @ [0, h, 0]
        np.testing.assert_allclose(got[i, :3], np.minimum(a, b) - rad, atol=2e-6 * (1 + np.abs(t).max()))
        np.testing.assert_allclose(got[i, 3:], np.maximum(a, b) + rad, atol=2e-6 * (1 + np.abs(t).max()))


def cone_world_aabb(R, t, hh, r):
    """The cone is the hull of its apex t + hh u and the rim circle around
    c = t - hh u of radius r in the plane normal to u = R e_y.  Along a world axis
    the circle reaches c_d +- r sqrt(1 - u_d^2)."""
    u = R[:, 1]
    apex, c = t + hh * u, t - hh * u
    ext = r * np.sqrt(np.maximum(0.0, 1.0 - u * u))
    return np.minimum(apex, c - ext), np.maximum(apex, c + ext)


def test_cone_aabb_closed_form(orc):
    rng = np.random.default_rng(23)
    n = 400
    hh = rng.uniform(0.3, 1.5, n)
    r = rng.uniform(0.3, 1.5, n)
    descs = np.zeros(n, SHAPE_DESC_DTYPE)
    descs["type"] = SH_CONE
    descs["params"][:, 0], descs["params"][:, 1] = hh, r
    poses = _random_poses(rng, n)
    # axis-aligned and upside-down cones reach the rim's zero-length branch of the formula
    poses[:3, :4] = [[0, 0, 0, 1], [1, 0, 0, 0], [0, 0, np.sqrt(0.5), np.sqrt(0.5)]]
    got = orc.shape_aabbs(descs, np.zeros((0, 3), np.float32), poses).astype(np.float64)
    for i in range(n):
        R = _rot(poses[i, :4])
        t = poses[i, 4:].astype(np.float64)
        lo, hi = cone_world_aabb(R, t, np.float64(np.float32(hh[i])), np.float64(np.float32(r[i])))
        np.testing.assert_allclose(got[i, :3], lo, atol=3e-6 * (1 + np.abs(t).max()))
        np.testing.assert_allclose(got[i, 3:], hi, atol=3e-6 * (1 + np.abs(t).max()))


def test_ball_aabb_closed_form(orc):
    rng = np.random.default_rng(24)
    n = 50
    descs = np.zeros(n, SHAPE_DESC_DTYPE)
    descs["type"] = SH_BALL
    descs["params"][:, 0] = rng.uniform(0.2, 2.0, n)
    poses = _random_poses(rng, n)
    got = orc.shape_aabbs(descs, np.zeros((0, 3), np.float32), poses)
    r = descs["params"][:, 0][:, None]
    assert np.array_equal(got[:, :3], poses[:, 4:] - r) and np.array_equal(got[:, 3:], poses[:, 4:] + r)


# ---- support maps ------------------------------------------------------------------


def _directions(rng, n):
    d = rng.standard_normal((n, 3))
    special = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, -1], [1, 1e-6, 0], [1e-6, 1, 0], [0.3, -2.0, 0.1],
                        [2.0, 0.0, -1.0], [-1e-3, 1.0, 1e-3], [5.0, -0.2, 0.0]])
    return np.concatenate([special, d])


def _dense_capsule_surface(hh, r, m=400):
    """Points on the capsule surface: the two hemispheres and the cylinder."""
    g = np.random.default_rng(0).standard_normal((m * m // 4, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    top = g * r + [0, hh, 0]
    bot = g * r - [0, hh, 0]
    th = np.linspace(0, 2 * np.pi, m, endpoint=False)
    ys = np.linspace(-hh, hh, 9)
    cyl = np.array([[r * np.cos(a), y, r * np.sin(a)] for a in th for y in ys])
    return np.concatenate([top, bot, cyl])


def _dense_cone_surface(hh, r, m=20000):
    """Points on the cone surface: apex, rim, lateral lines and the base disc."""
    th = np.linspace(0, 2 * np.pi, m, endpoint=False)
    rim = np.stack([r * np.cos(th), np.full(m, -hh), r * np.sin(th)], 1)
    s = np.random.default_rng(1).uniform(0, 1, (m, 1))
    lateral = (1 - s) * rim + s * np.array([0, hh, 0])
    base = rim * np.random.default_rng(2).uniform(0, 1, (m, 1))
    base[:, 1] = -hh
    return np.concatenate([[[0, hh, 0]], rim, lateral, base])


@pytest.mark.parametrize("f64", [False, True])
def test_capsule_support_closed_form_and_samples(orc, f64):
    rng = np.random.default_rng(31)
    tol = 1e-12 if f64 else 2e-6
    for hh, r in ((0.7, 0.3), (1.3, 0.55), (0.5, 0.75)):
        hh, r = float(np.float32(hh)), float(np.float32(r))
        dirs = _directions(rng, 300)
        p = orc.local_support_points(desc(SH_CAPSULE, params=(hh, r, 0, 0)), None, dirs, f64=f64).astype(np.float64)
        dn = np.linalg.norm(dirs, axis=1)
        h = hh * np.abs(dirs[:, 1]) + r * dn  # support function of the capsule
        assert (np.abs(np.einsum("ij,ij->i", p, dirs) - h) <= tol * 10 * (1 + dn)).all()
        # the point lies on the surface: distance r from the segment (0, -hh, 0)-(0, hh, 0)
        seg = np.clip(p[:, 1], -hh, hh)
        dist = np.sqrt(p[:, 0] ** 2 + (p[:, 1] - seg) ** 2 + p[:, 2] ** 2)
        np.testing.assert_allclose(dist, r, atol=tol * 10)
        # no sampled surface point does better
        surf = _dense_capsule_surface(hh, r)
        best = (surf @ dirs.T).max(0)
        assert (np.einsum("ij,ij->i", p, dirs) >= best - tol * 10 * (1 + dn)).all()
    # zero direction: Unit::try_new(0, 0) fails and +y is used (capsule.rs:408-410)
    p0 = orc.local_support_points(desc(SH_CAPSULE, params=(0.7, 0.3, 0, 0)), None, np.zeros((1, 3)), f64=f64)
    np.testing.assert_allclose(p0[0], [0, np.float32(0.7) + np.float32(0.3), 0], rtol=1e-6)


@pytest.mark.parametrize("f64", [False, True])
def test_cone_support_closed_form_and_samples(orc, f64):
    rng = np.random.default_rng(32)
    tol = 1e-12 if f64 else 2e-6
    for hh, r in ((0.7, 0.3), (1.3, 1.1), (0.5, 1.45)):
        hh, r = float(np.float32(hh)), float(np.float32(r))
        dirs = _directions(rng, 300)
        p = orc.local_support_points(desc(SH_CONE, params=(hh, r, 0, 0)), None, dirs, f64=f64).astype(np.float64)
        dn = np.linalg.norm(dirs, axis=1)
        dxz = np.hypot(dirs[:, 0], dirs[:, 2])
        h = np.maximum(hh * dirs[:, 1], -hh * dirs[:, 1] + r * dxz)  # support function of the cone
        assert (np.abs(np.einsum("ij,ij->i", p, dirs) - h) <= tol * 10 * (1 + dn)).all()
        # the point is the apex, the rim, or (d parallel to -y) the base centre
        apex = np.all(np.abs(p - [0, hh, 0]) <= tol, axis=1)
        rim = (np.abs(p[:, 1] + hh) <= tol) & (np.abs(np.hypot(p[:, 0], p[:, 2]) - r) <= tol * 10)
        centre = np.all(np.abs(p - [0, -hh, 0]) <= tol, axis=1) & (dxz == 0)
        assert (apex | rim | centre).all()
        surf = _dense_cone_surface(hh, r)
        best = (surf @ dirs.T).max(0)
        assert (np.einsum("ij,ij->i", p, dirs) >= best - tol * 10 * (1 + dn)).all()
        # rim points lie along the direction's horizontal part
        sel = rim & (dxz > 1e-3)
        np.testing.assert_allclose(p[sel][:, [0, 2]] / r, dirs[sel][:, [0, 2]] / dxz[sel, None], atol=tol * 100)
    # axial known answers (cone.rs:166-168): +y gives the apex, -y the base centre
    d = desc(SH_CONE, params=(0.7, 0.4, 0, 0))
    p = orc.local_support_points(d, None, np.array([[0, 2.0, 0], [0, -3.0, 0]]), f64=f64)
    np.testing.assert_allclose(p, [[0, np.float32(0.7), 0], [0, -np.float32(0.7), 0]], rtol=1e-7)


@pytest.mark.parametrize("f64", [False, True])
def test_cuboid_and_ball_support(orc, f64):
    rng = np.random.default_rng(33)
    dirs = _directions(rng, 200)
    he = np.array([0.4, 1.1, 0.7])
    p = orc.local_support_points(desc(SH_CUBOID, params=(*he, 0)), None, dirs, f64=f64).astype(np.float64)
    # the corner whose signs are those of the direction (a zero component takes its sign bit)
    exp = np.where(np.signbit(dirs), -1.0, 1.0) * np.float32(he).astype(np.float64)
    assert np.array_equal(p, exp)
    p = orc.local_support_points(desc(SH_BALL, params=(0.8, 0, 0, 0)), None, dirs, f64=f64).astype(np.float64)
    np.testing.assert_allclose(p, dirs / np.linalg.norm(dirs, axis=1, keepdims=True) * np.float32(0.8),
                               atol=1e-12 if f64 else 1e-6)


# ---- Capsule-Capsule contact against the segment distance -------------------------


def _seg_point(p, a, b):
    """Closest point to p on segment ab (float64)."""
    ab = b - a
    s = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return a + s * ab


def segment_segment(a0, a1, b0, b1):
    """Closest points of two segments by a ternary search over the first
    segment's parameter (the distance to a convex set is convex along a line)."""
    lo, hi = 0.0, 1.0
    f = lambda s: np.linalg.norm(a0 + s * (a1 - a0) - _seg_point(a0 + s * (a1 - a0), b0, b1))
    for _ in range(200):
        m1, m2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
        if f(m1) <= f(m2):
            hi = m2
        else:
            lo = m1
    s = 0.5 * (lo + hi)
    pa = a0 + s * (a1 - a0)
    return pa, _seg_point(pa, b0, b1)


def _capsule_gap(ends, radii, n):
    """min over capsule 2 of n.y minus max over capsule 1 of n.x, for unit n."""
    (a0, a1), (b0, b1) = ends
    hi1 = max(np.dot(n, a0), np.dot(n, a1)) + radii[0]
    lo2 = min(np.dot(n, b0), np.dot(n, b1)) - radii[1]
    return lo2 - hi1


def test_capsule_capsule_contact_segment_distance(orc):
    rng = np.random.default_rng(41)
    n = 300
    hh = rng.uniform(0.4, 1.5, 2 * n)
    rr = rng.uniform(0.2, 0.7, 2 * n)
    descs = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    descs["type"] = SH_CAPSULE
    descs["params"][:, 0], descs["params"][:, 1] = hh, rr
    poses = _random_poses(rng, 2 * n, spread=1.6)
    pairs = np.arange(2 * n, dtype=np.uint32).reshape(-1, 2)
    pred = 0.5
    r64 = orc.contact_pairs(descs, np.zeros((0, 3)), poses.astype(np.float64), pairs, pred, f64=True)
    r32 = orc.contact_pairs(descs, np.zeros((0, 3), np.float32), poses, pairs, pred)
    P = poses.astype(np.float64)
    n_sep = n_pen = 0
    for k in range(n):
        i, j = 2 * k, 2 * k + 1
        ends = []
        for s in (i, j):
            R, h = _rot(P[s, :4]), np.float64(descs["params"][s, 0])
            ends.append((P[s, 4:] + R @ [0, -h, 0], P[s, 4:] + R @ [0, h, 0]))
        pa, pb = segment_segment(*ends[0], *ends[1])
        ra, rb = np.float64(descs["params"][i, 1]), np.float64(descs["params"][j, 1])
        dist = np.linalg.norm(pb - pa) - ra - rb
        if dist > pred + 1e-6:
            assert r64["status"][k] == 0 and r32["status"][k] == 0
            continue
        if dist >= pred - 1e-6:
            continue  # on the prediction boundary
        assert r64["status"][k] == 1 and r32["status"][k] == 1
        if dist > 1e-3:
            # separated.  GJK's f64 stopping rule is relative (sqrt(10 EPS) ~ 5e-8), but on
            # curved shapes GJK can also stop when its upper bound stalls, a few 1e-6 short;
            # the f32-rounded quaternions bend lengths by ~1e-7.
            # In f32 the stall comes earlier on these curved shapes (up to ~1% here).
            assert abs(r64["out"][k, 12] - dist) <= 1e-5
            assert abs(r32["out"][k, 12] - dist) <= 1e-2 * max(1.0, dist)
            nrm = (pb - pa) / np.linalg.norm(pb - pa)
            # normal1 separates the capsules by dist (their support functions along it);
            # near-parallel segments make its direction ill-conditioned, not this gap
            assert abs(_capsule_gap(ends, (ra, rb), r64["out"][k, 6:9]) - dist) <= 1e-5
            assert np.dot(r64["out"][k, 6:9], nrm) > 1 - 1e-4
            # witness points: on each capsule's surface
            for w, (e0, e1), rad in ((r64["out"][k, 0:3], ends[0], ra), (r64["out"][k, 3:6], ends[1], rb)):
                assert abs(np.linalg.norm(w - _seg_point(w, e0, e1)) - rad) <= 1e-5
            n_sep += 1
        elif dist < -1e-3 and np.linalg.norm(pb - pa) > 1e-2:
            # penetrating with distinct segment points: EPA on the rounded shapes
            # approximates the curved Minkowski difference by a polytope
            assert abs(r64["out"][k, 12] - dist) <= 1e-5 * max(1.0, -dist)
            assert abs(r32["out"][k, 12] - dist) <= 1e-2 * max(1.0, -dist)
            assert abs(_capsule_gap(ends, (ra, rb), r64["out"][k, 6:9]) - dist) <= 1e-5 * max(1.0, -dist)
            n_pen += 1
    assert n_sep >= 30 and n_pen >= 30, (n_sep, n_pen)


# ---- Cone distances against constrained minimisation ------------------------------


def _dist_to_cone(p, Rc, tc, hh, r):
    """Distance from a world point to the posed cone, in the cone's (rho, y)
    half-plane: the cone's cross-section there is the triangle (0, hh),
    (r, -hh), (0, -hh), and rotation about the axis keeps distances."""
    x = Rc.T @ (p - tc)
    q = np.array([np.hypot(x[0], x[2]), x[1]])
    tri = np.array([[0.0, hh], [r, -hh], [0.0, -hh]])
    inside = q[1] >= -hh and q[1] <= hh and q[0] <= r * (hh - q[1]) / (2 * hh)
    if inside:
        return 0.0
    best = np.inf
    for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
        ab = b - a
        s = np.clip(np.dot(q - a, ab) / np.dot(ab, ab), 0.0, 1.0)
        best = min(best, np.linalg.norm(q - (a + s * ab)))
    return best


def _cone_dist_slsqp(Rc, tc, hh, r, halfspaces):
    """min over y in {y : A y + b <= 0} of the distance from y to the posed cone:
    a convex function (C^1 outside the cone) over a polytope, minimised by SLSQP
    (float64) from a few starts."""
    from scipy.optimize import minimize

    A, b, inner = halfspaces
    cons = [{"type": "ineq", "fun": lambda y: -(A @ y + b), "jac": lambda y: -A}]
    best = np.inf
    for start in range(3):
        y0 = inner + np.random.default_rng(start).normal(0, 0.05, 3)
        res = minimize(lambda y: _dist_to_cone(y, Rc, tc, hh, r), y0, constraints=cons, method="SLSQP",
                       options={"ftol": 1e-14, "maxiter": 1000})
        if np.max(A @ res.x + b) <= 1e-9:
            best = min(best, res.fun)
    assert np.isfinite(best)
    return best


def _box_halfspaces(R, t, he):
    n = np.concatenate([R.T, -R.T])  # world normals of the six faces
    b = -(n @ t) - np.concatenate([he, he])
    return n, b, t


def _hull_halfspaces(Pw):
    from scipy.spatial import ConvexHull

    eq = ConvexHull(Pw).equations
    return eq[:, :3], eq[:, 3], Pw.mean(0)


def test_cone_cuboid_and_polyhedron_distance(orc):
    rng = np.random.default_rng(51)
    n = 40
    verts, offs = scenes.convex_polyhedra(n, 52, kmin=8, kmax=24)
    descs = np.zeros(3 * n, SHAPE_DESC_DTYPE)
    descs["type"][0::3] = SH_CONE
    descs["params"][0::3, 0] = rng.uniform(0.4, 1.2, n)
    descs["params"][0::3, 1] = rng.uniform(0.3, 1.2, n)
    descs["type"][1::3] = SH_CUBOID
    descs["params"][1::3, :3] = rng.uniform(0.3, 1.0, (n, 3))
    descs["type"][2::3] = SH_POLY
    descs["first_vertex"][2::3] = offs[:-1]
    descs["num_vertices"][2::3] = offs[1:] - offs[:-1]
    poses = np.zeros((3 * n, 7), np.float32)
    q = scenes.random_rotations(3 * n, rng)
    poses[:, :4] = q
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    poses[1::3, 4:] = d * rng.uniform(2.6, 3.4, (n, 1))
    poses[2::3, 4:] = -d * rng.uniform(2.8, 3.6, (n, 1))
    pairs = np.array([[3 * k, 3 * k + s] for k in range(n) for s in (1, 2)], np.uint32)
    r64 = orc.contact_pairs(descs, verts.astype(np.float64), poses.astype(np.float64), pairs, 10.0, f64=True)
    r32 = orc.contact_pairs(descs, verts, poses, pairs, 10.0)
    P = poses.astype(np.float64)
    checked = 0
    for k, (i, j) in enumerate(pairs):
        Rc, tc = _rot(P[i, :4]), P[i, 4:]
        hh, r = np.float64(descs["params"][i, 0]), np.float64(descs["params"][i, 1])
        R2, t2 = _rot(P[j, :4]), P[j, 4:]
        if descs["type"][j] == SH_CUBOID:
            hs = _box_halfspaces(R2, t2, descs["params"][j, :3].astype(np.float64))
        else:
            V = verts[descs["first_vertex"][j]:descs["first_vertex"][j] + descs["num_vertices"][j]].astype(np.float64)
            hs = _hull_halfspaces(V @ R2.T + t2)
        dref = _cone_dist_slsqp(Rc, tc, hh, r, hs)
        if dref < 0.05:
            continue
        assert r64["status"][k] == 1 and r32["status"][k] == 1
        # GJK's witness points lie on the shapes, so its distance is an upper bound that
        # stops within its tolerance (a few 1e-6 on the cone's curved rim in f64)
        # (the f32 quaternions are unit only to ~1e-7, which bends lengths that much)
        assert -1e-6 <= r64["out"][k, 12] - dref <= 1e-5 * max(1.0, dref), (k, r64["out"][k, 12], dref)
        assert abs(r32["out"][k, 12] - dref) <= 1e-2 * max(1.0, dref)
        checked += 1
    assert checked >= 60


# ---- hull vertex order (ConvexPolyhedron::from_convex_hull) --------------------------


def test_remove_unused_points_order_known_answers():
    """remove_unused_points (cleanup.rs:6-34) traced by hand: an unused point is
    replaced by the last point (Vec::swap_remove) and its slot examined again."""
    from parry_amd.scenes import remove_unused_points_order as order

    # [p0..p5], used {2, 3, 5}: swap_remove(0) -> [p5 p1 p2 p3 p4]; p5 used, i=1;
    # swap_remove(1) -> [p5 p4 p2 p3]; p4 unused: swap_remove(1) -> [p5 p3 p2]
    assert order(6, [0, 0, 1, 1, 0, 1]) == [5, 3, 2]
    assert order(5, [1, 0, 1, 0, 1]) == [0, 4, 2]
    assert order(4, [1, 1, 1, 1]) == [0, 1, 2, 3]
    assert order(4, [0, 1, 1, 1]) == [3, 1, 2]
    assert order(3, [1, 1, 0]) == [0, 1]  # the last point just goes
    assert order(3, [0, 0, 0]) == []


def test_hull_vertices_keep_triangles_pointing_at_the_same_points():
    """With the remap remove_unused_points builds, every hull triangle still names
    the same coordinates; every hull vertex is kept exactly once."""
    from scipy.spatial import ConvexHull

    from parry_amd.scenes import convex_hull_vertices, convex_point_cloud

    rng = np.random.default_rng(61)
    for _ in range(50):
        pts = convex_point_cloud(rng, int(rng.integers(8, 65)))
        h = ConvexHull(pts)
        v = convex_hull_vertices(pts)
        assert v.shape[0] == len(h.vertices)
        pos = {tuple(p): k for k, p in enumerate(v)}
        assert len(pos) == v.shape[0]
        for tri in h.simplices:
            for i in tri:
                assert np.array_equal(v[pos[tuple(pts[i])]], pts[i])
