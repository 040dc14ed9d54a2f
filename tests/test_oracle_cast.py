"""Oracle pins for query::cast_shapes (shape_cast.rs:274-286): linear shape casts.

  * the reference's own tests: ball_cuboid_toi (time_of_impact3.rs), test_ball_ball_toi
    (ball_ball_toi.rs, exactly 0.9), ball_triangle_toi_infinite_loop_issue
    (ball_triangle_toi.rs, None) and the stacked cuboids of still_objects_toi.rs;
  * Ball-Ball (closed form) against an independent float64 solution of
    |c + v t| = r1 + r2;
  * polytope casts (ConvexPolyhedron, Cuboid, Triangle; GJK's ray cast on the
    Minkowski difference) against the exact time of impact of a linear program:
    min t such that sum(l_i a_i) - sum(m_j b_j) - t v = 0 with l and m on the
    unit simplices -- independent of GJK;
  * target_distance on a Ball equals the cast of a ball grown by it;
  * the penetration branch: toi 0 with the contact's normal and points, and None
    with stop_at_penetration = false for separating motion;
  * f32 against f64.
"""
import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import oracle as orc
from parry_amd import scenes
from parry_amd.shape import SHAPE_DESC_DTYPE

IDQ = [0.0, 0.0, 0.0, 1.0]


def _descs(types, params, nverts=None, first=None):
    d = np.zeros(len(types), SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, p in enumerate(params):
        d["params"][i][: len(p)] = p
    if nverts is not None:
        d["num_vertices"] = nverts
        d["first_vertex"] = first
    return d


def _cast1(descs, verts, p1, v1, p2, v2, **kw):
    poses = np.array([p1, p2], np.float32)
    vels = np.array([v1, v2], np.float32)
    return orc.cast_shapes(descs, verts, poses, vels, np.array([[0, 1]], np.uint32), **kw)


NOV = np.zeros((0, 3), np.float32)


def test_reference_ball_cuboid_toi():
    # time_of_impact3.rs: ball (r=1) first, cuboid (he=1) second
    d = _descs([0, 1], [[1.0], [1.0, 1.0, 1.0]])
    r = _cast1(d, NOV, IDQ + [1, 1, 1], [2, 2, 2], IDQ + [0, 0, 0], [-1, 1, 1])
    assert r["status"][0] != 0 and r["out"][0, 0] == 0.0
    r = _cast1(d, NOV, IDQ + [2, 2, 2], [-0.5, -0.5, -0.5], IDQ + [0, 0, 0], [1, 1, 1])
    expect = (np.float32(np.sqrt(np.float32(3.0))) - np.float32(1.0)) / np.linalg.norm(
        np.array([-1.5, -1.5, -1.5], np.float32))
    toi = r["out"][0, 0]
    assert r["status"][0] == orc.CAST_CONVERGED
    assert abs(toi - expect) <= max(abs(toi), abs(expect)) * np.finfo(np.float32).eps or \
        abs(toi - expect) <= np.finfo(np.float32).eps  # relative_eq!
    r = _cast1(d, NOV, IDQ + [3, 3, 3], [2, 2, 2], IDQ + [0, 0, 0], [-1, 1, 1])
    assert r["status"][0] == 0


def test_reference_ball_ball_toi():
    d = _descs([0, 0], [[0.5], [0.5]])
    r = _cast1(d, NOV, IDQ + [0, 0, 0], [0, 10, 0], IDQ + [0, 10, 0], [0, 0, 0])
    assert r["out"][0, 0] == np.float32(0.9)


def test_reference_ball_triangle_tiny_velocity():
    tri = np.array([[0.5, -0.5, 0.0], [-0.5, -0.5, 0.0], [-0.5, 0.5, 0.0]], np.float32)
    d = _descs([0, 5], [[0.375], []], nverts=[0, 3], first=[0, 0])
    r = _cast1(d, tri, IDQ + [0, 0, 0], [0, 6.925e-42, 0], IDQ + [11.5, 5.5, 0], [0, 0, 0])
    assert r["status"][0] == 0


@pytest.mark.parametrize("vy,hit", [(0.0, False), (1.0, False), (-1.0, True)])
def test_reference_still_objects(vy, hit):
    d = _descs([1, 1], [[0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    r = _cast1(d, NOV, IDQ + [0, 1.1, 0], [0, vy, 0], IDQ + [0, 0, 0], [0, 0, 0])
    assert (r["status"][0] != 0) == hit
    if hit:  # the faces 0.1 apart close at unit speed
        assert abs(r["out"][0, 0] - 0.1) < 1e-5


def _rotate(q, v):
    q = np.asarray(q, np.float64)
    u, w = q[:3], q[3]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def test_ball_ball_closed_form():
    rng = np.random.default_rng(3)
    n = 2000
    d = _descs([0] * (2 * n), [[r] for r in rng.uniform(0.1, 1.0, 2 * n)])
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), rng.uniform(-3, 3, (2 * n, 3))], 1).astype(np.float32)
    vels = rng.uniform(-2, 2, (2 * n, 3)).astype(np.float32)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    r = orc.cast_shapes(d, NOV, poses, vels, pairs)
    hits = 0
    for k in range(n):
        i, j = pairs[k]
        c = poses[j, 4:].astype(np.float64) - poses[i, 4:]
        v = vels[j].astype(np.float64) - vels[i]
        rs = float(d["params"][i][0]) + float(d["params"][j][0])
        # |c + v t| = rs, smallest t >= 0 (0 if already overlapping)
        a, b, cc = v @ v, c @ v, c @ c - rs * rs
        if cc <= 0:
            ref = 0.0
        else:
            disc = b * b - a * cc
            ref = None if (b >= 0 or disc < 0) else (-b - np.sqrt(disc)) / a
        if ref is None:
            assert r["status"][k] == 0
        else:
            hits += 1
            assert r["status"][k] != 0
            assert abs(r["out"][k, 0] - ref) <= 2e-5 * max(1.0, ref)
            # normal1 is the unit direction from ball 1's centre to the contact, in ball 1's frame
            n1 = _rotate(poses[i, :4] * [-1, -1, -1, 1], c + v * ref)
            if cc > 0:
                np.testing.assert_allclose(r["out"][k, 7:10], n1 / np.linalg.norm(n1), atol=1e-4)
    assert hits > 80


def _lp_toi(a, b, v):
    """Exact time of impact of polytope B (vertices b, moving by v) onto polytope A
    (vertices a), both in A's frame: min t s.t. sum l a - sum m b - t v = 0."""
    na_, nb = len(a), len(b)
    c = np.zeros(na_ + nb + 1)
    c[-1] = 1.0
    aeq = np.zeros((5, na_ + nb + 1))
    aeq[:3, :na_] = a.T
    aeq[:3, na_:na_ + nb] = -b.T
    aeq[:3, -1] = -v
    aeq[3, :na_] = 1.0
    aeq[4, na_:na_ + nb] = 1.0
    beq = np.array([0, 0, 0, 1, 1], np.float64)
    res = linprog(c, A_eq=aeq, b_eq=beq, bounds=[(0, None)] * (na_ + nb + 1), method="highs")
    return res.x[-1] if res.status == 0 else None


def _posed(pose, pts):
    return np.array([_rotate(pose[:4], p) + pose[4:] for p in pts.astype(np.float64)])


def _cuboid_pts(he):
    s = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    return s * he


def test_polytope_casts_against_linear_program():
    rng = np.random.default_rng(11)
    pv, offs = scenes.convex_polyhedra(40, seed=4)
    pv = pv.astype(np.float32)
    n = 120
    types = rng.choice([1, 4, 5], size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = types
    verts = [pv]
    nv = pv.shape[0]
    pts = []
    for s in range(2 * n):
        if types[s] == 1:
            he = rng.uniform(0.3, 1.0, 3)
            d["params"][s][:3] = he
            pts.append(_cuboid_pts(he.astype(np.float32).astype(np.float64)))
        elif types[s] == 4:
            k = int(rng.integers(0, 40))
            d["first_vertex"][s], d["num_vertices"][s] = offs[k], offs[k + 1] - offs[k]
            pts.append(pv[offs[k]:offs[k + 1]].astype(np.float64))
        else:
            t = rng.normal(0, 0.8, (3, 3)).astype(np.float32)
            d["first_vertex"][s], d["num_vertices"][s] = nv, 3
            verts.append(t)
            nv += 3
            pts.append(t.astype(np.float64))
    verts = np.concatenate(verts, 0)
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    # shape 2 starts 3-5 units away and moves roughly towards shape 1
    dirs = rng.normal(0, 1, (n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    poses[1::2, 4:] = dirs * rng.uniform(3.0, 5.0, (n, 1))
    poses = poses.astype(np.float32)
    vels = np.zeros((2 * n, 3), np.float32)
    vels[1::2] = (-dirs + rng.normal(0, 0.15, (n, 3))) * rng.uniform(0.5, 2.0, (n, 1))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    r = orc.cast_shapes(d, verts, poses, vels, pairs)
    hits = 0
    for k in range(n):
        i, j = pairs[k]
        qi = poses[i, :4].astype(np.float64) * [-1, -1, -1, 1]
        a = pts[i]
        b = np.array([_rotate(qi, _rotate(poses[j, :4], p) + poses[j, 4:] - poses[i, 4:]) for p in pts[j]])
        v = _rotate(qi, vels[j].astype(np.float64) - vels[i])
        ref = _lp_toi(a, b, v)
        if ref is None:
            assert r["status"][k] == 0, k
        else:
            hits += 1
            assert r["status"][k] == orc.CAST_CONVERGED, k
            assert abs(r["out"][k, 0] - ref) <= 1e-4 * max(1.0, ref), (k, r["out"][k, 0], ref)
    assert hits > 40 and hits < n


def test_target_distance_is_a_grown_ball():
    rng = np.random.default_rng(8)
    n = 300
    # Ball (r) first, Cuboid second; target distance t against Ball (r + t)
    r0 = rng.uniform(0.2, 0.6, n)
    td = rng.uniform(0.05, 0.3, n)
    he = rng.uniform(0.3, 1.0, (n, 3))
    d1 = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d2 = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d1["type"][0::2] = 0
    d1["type"][1::2] = 1
    d2["type"][:] = d1["type"]
    d1["params"][0::2, 0] = r0
    d2["params"][0::2, 0] = r0 + td
    d1["params"][1::2, :3] = he
    d2["params"][1::2, :3] = he
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    dirs = rng.normal(0, 1, (n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    poses[1::2, 4:] = dirs * 3.0
    vels = np.zeros((2 * n, 3))
    vels[1::2] = -dirs + rng.normal(0, 0.2, (n, 3))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    got = [orc.cast_shapes(d1, NOV, poses, vels, pairs[k:k + 1], target_distance=float(td[k])) for k in range(n)]
    ref = orc.cast_shapes(d2, NOV, poses, vels, pairs)
    st = np.array([g["status"][0] for g in got])
    toi = np.array([g["out"][0, 0] for g in got])
    assert np.array_equal(st != 0, ref["status"] != 0)
    h = st != 0
    assert h.sum() > 50
    np.testing.assert_allclose(toi[h], ref["out"][h, 0], atol=2e-4, rtol=2e-4)


def test_penetration_branch():
    # two overlapping cuboids: toi 0 and the contact's geometry (dist < 0)
    d = _descs([1, 1], [[0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    poses = np.array([IDQ + [0, 0, 0], IDQ + [0.8, 0.1, 0]], np.float32)
    vels = np.array([[0, 0, 0], [-1, 0, 0]], np.float32)
    pairs = np.array([[0, 1]], np.uint32)
    r = orc.cast_shapes(d, NOV, poses, vels, pairs)
    assert r["status"][0] == orc.CAST_PENETRATING and r["out"][0, 0] == 0.0
    c = orc.contact_pairs(d, NOV, poses, pairs, 1e30)
    np.testing.assert_allclose(r["out"][0, 7:10], c["out"][0, 6:9], atol=1e-6)  # normal1 = +x (local = world here)
    np.testing.assert_allclose(r["out"][0, 7:10], [1, 0, 0], atol=1e-6)
    # moving apart: None when stop_at_penetration is false, a hit when true
    sep = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    assert orc.cast_shapes(d, NOV, poses, sep, pairs, stop_at_penetration=False)["status"][0] == 0
    assert orc.cast_shapes(d, NOV, poses, sep, pairs)["status"][0] == orc.CAST_PENETRATING
    # without impact geometry: toi 0, status penetrating, the ray cast's own values
    r2 = orc.cast_shapes(d, NOV, poses, vels, pairs, compute_impact_geometry_on_penetration=False)
    assert r2["status"][0] == orc.CAST_PENETRATING and r2["out"][0, 0] == 0.0


def test_f32_agrees_with_f64():
    rng = np.random.default_rng(21)
    descs, verts = scenes.mixed_shapes(800, 5)
    n = 400
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    dirs = rng.normal(0, 1, (n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    poses[1::2, 4:] = dirs * rng.uniform(2.0, 6.0, (n, 1))
    vels = np.zeros((2 * n, 3))
    vels[1::2] = (-dirs + rng.normal(0, 0.3, (n, 3))) * rng.uniform(0.5, 3.0, (n, 1))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    r32 = orc.cast_shapes(descs, verts, poses, vels, pairs)
    r64 = orc.cast_shapes(descs, verts, poses, vels, pairs, f64=True)
    both = (r32["status"] != 0) & (r64["status"] != 0)
    assert (r32["status"] == r64["status"]).mean() > 0.97
    assert both.sum() > 100
    d = np.abs(r32["out"][both, 0] - r64["out"][both, 0])
    assert np.quantile(d, 0.99) < 1e-3
