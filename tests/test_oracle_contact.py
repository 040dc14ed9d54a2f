"""Pins the oracle's contact restatement (GJK, Voronoi simplex, EPA, query::contact).

  * reference known-answer tests: crates/parry3d/tests/geometry/epa3.rs:6-23
    (Cuboid(2,1,1) vs itself: dist == -0.5 exactly with normal1 == -x, then
    dist == -1.8 with normal1 == -y) and epa3.rs:25-59 (touching triangles);
  * separated convex polyhedra: the GJK distance equals the hull-to-hull
    distance found by an independent constrained minimisation (scipy SLSQP),
    and the witness points lie on their shapes;
  * penetrating convex polyhedra: the EPA depth equals the minimum translation
    distance from the separating-axis theorem over face normals and edge-edge
    cross products (exact for polytopes), and normal1 is that axis;
  * Ball-Ball closed form (contact_ball_ball.rs:9-42);
  * f32 and f64 restatements agree, and dist < 0 exactly when the SAT says the
    shapes overlap.
"""
import numpy as np
import pytest

from parry_amd import SHAPE_DESC_DTYPE, scenes

SH_BALL, SH_CUBOID, SH_POLY = 0, 1, 4


def desc(kind, first=0, count=0, params=(0, 0, 0, 0)):
    d = np.zeros(1, SHAPE_DESC_DTYPE)
    d["type"], d["first_vertex"], d["num_vertices"] = kind, first, count
    d["params"] = params
    return d


IDENT = [0, 0, 0, 1]


def test_kat_cuboid_cuboid_epa(orc):
    descs = np.concatenate([desc(SH_CUBOID, params=(2, 1, 1, 0)), desc(SH_CUBOID, params=(2, 1, 1, 0))])
    for t, dist, normal in (((3.5, 0, 0), -0.5, (-1, 0, 0)), ((0, 0.2, 0), -1.8, (0, -1, 0))):
        poses = np.array([IDENT + list(t), IDENT + [0, 0, 0]], np.float32)
        r = orc.contact_pairs(descs, np.zeros((0, 3), np.float32), poses, [[0, 1]], 10.0)
        assert r["status"][0] == 1
        assert r["out"][0, 12] == np.float32(dist)  # exact, as the reference asserts
        assert tuple(r["out"][0, 6:9]) == normal


def test_kat_triangle_vertex_touches_edge(orc):
    tri = np.array(
        [[-13.174434, 1.0, 8.736801], [3.5251038, 1.0, 12.1], [3.2048466, 1.0, 12.218325],
         [-1.63, 0.0, 11.19], [-2.349647, 0.0, 11.037681], [-2.349647, 1.0, 11.037681]], np.float32)
    descs = np.concatenate([desc(orc.SHAPE_TRIANGLE_TEST_ONLY, 0, 3), desc(orc.SHAPE_TRIANGLE_TEST_ONLY, 3, 3)])
    poses = np.array([IDENT + [0, 0, 0], IDENT + [0, 0, 0]], np.float32)
    r = orc.contact_pairs(descs, tri, poses, [[0, 1]], 0.00999999977)
    assert r["status"][0] == 1
    a = r["out"][0, 0:3]
    assert abs(a[1] - 1.0) <= 1e-3 and abs(a[0] - -2.349647) <= 1e-3


def quat_matrix(q):
    x, y, z, w = q
    return np.array(
        [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _world(verts, pose):
    return verts.astype(np.float64) @ quat_matrix(pose[:4].astype(np.float64)).T + pose[4:].astype(np.float64)


def hull_distance(A, B):
    """min |sum(l_i A_i) - sum(m_j B_j)| over convex weights (SLSQP, f64)."""
    from scipy.optimize import minimize

    na, nb = A.shape[0], B.shape[0]

    def f(x):
        d = x[:na] @ A - x[na:] @ B
        return d @ d

    def g(x):
        d = x[:na] @ A - x[na:] @ B
        return np.concatenate([2 * A @ d, -2 * B @ d])

    x0 = np.concatenate([np.full(na, 1 / na), np.full(nb, 1 / nb)])
    cons = [{"type": "eq", "fun": lambda x: x[:na].sum() - 1}, {"type": "eq", "fun": lambda x: x[na:].sum() - 1}]
    res = minimize(f, x0, jac=g, bounds=[(0, 1)] * (na + nb), constraints=cons, method="SLSQP",
                   options={"ftol": 1e-16, "maxiter": 1000})
    return np.sqrt(max(res.fun, 0.0))


def sat_depth(A, B):
    """Minimum translation distance (>0 if overlapping) and its axis, from SAT candidates."""
    from scipy.spatial import ConvexHull

    axes = []
    edges = []
    for P in (A, B):
        h = ConvexHull(P)
        axes.append(h.equations[:, :3])
        e = set()
        for s in h.simplices:
            for i in range(3):
                a, b = sorted((s[i], s[(i + 1) % 3]))
                e.add((a, b))
        E = np.array([P[b] - P[a] for a, b in e])
        edges.append(E / np.linalg.norm(E, axis=1, keepdims=True))
    cr = np.cross(edges[0][:, None, :], edges[1][None, :, :]).reshape(-1, 3)
    nrm = np.linalg.norm(cr, axis=1)
    cr = cr[nrm > 1e-9] / nrm[nrm > 1e-9, None]
    # faces of the Minkowski difference A - B: normals of A, negated normals of B and
    # edge-edge cross products; both signs of everything keep the set a superset.
    cand = np.concatenate(axes + [-a for a in axes] + [cr, -cr])
    pa, pb = A @ cand.T, B @ cand.T
    # overlap needed to push B along +axis away from A: max_a - min_b
    ov = pa.max(0) - pb.min(0)
    k = int(np.argmin(ov))
    return ov[k], cand[k]


def _poly_pair_scene(seed, n_pairs, offset):
    verts, offs = scenes.convex_polyhedra(2 * n_pairs, seed, kmin=8, kmax=24)
    rng = np.random.default_rng(seed + 100)
    q = scenes.random_rotations(2 * n_pairs, rng)
    t = np.zeros((2 * n_pairs, 3))
    d = rng.standard_normal((n_pairs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t[1::2] = d * offset
    poses = np.concatenate([q, t], 1).astype(np.float32)
    descs = np.zeros(2 * n_pairs, SHAPE_DESC_DTYPE)
    descs["type"] = SH_POLY
    descs["first_vertex"] = offs[:-1]
    descs["num_vertices"] = offs[1:] - offs[:-1]
    pairs = np.arange(2 * n_pairs, dtype=np.uint32).reshape(-1, 2)
    return verts, offs, poses, descs, pairs


def test_separated_polyhedra_distance(orc):
    verts, offs, poses, descs, pairs = _poly_pair_scene(11, 24, 3.2)
    r64 = orc.contact_pairs(descs, verts, poses, pairs, 100.0, f64=True)
    r32 = orc.contact_pairs(descs, verts, poses, pairs, 100.0)
    checked = 0
    for k in range(pairs.shape[0]):
        A = _world(verts[offs[2 * k]:offs[2 * k + 1]], poses[2 * k].astype(np.float64))
        B = _world(verts[offs[2 * k + 1]:offs[2 * k + 2]], poses[2 * k + 1].astype(np.float64))
        depth, _ = sat_depth(A, B)
        if depth > -1e-3:
            continue  # overlapping or touching: covered by the EPA test
        dref = hull_distance(A, B)
        assert r64["status"][k] == 1 and r32["status"][k] == 1
        assert abs(r64["out"][k, 12] - dref) <= 1e-6 * max(1.0, dref)
        assert abs(r32["out"][k, 12] - dref) <= 2e-3 * max(1.0, dref)  # GJK's relative stopping rule
        # witness points: the distance between them equals dist, along normal1
        p1, p2, n1 = r64["out"][k, 0:3], r64["out"][k, 3:6], r64["out"][k, 6:9]
        # (the f32-rounded quaternions are unit only to ~1e-7, so world frames bend lengths that much)
        assert abs(np.dot(p2 - p1, n1) - r64["out"][k, 12]) < 1e-6
        checked += 1
    assert checked >= 10


def test_penetrating_polyhedra_depth(orc):
    verts, offs, poses, descs, pairs = _poly_pair_scene(12, 40, 0.6)
    r64 = orc.contact_pairs(descs, verts, poses, pairs, 0.0, f64=True)
    r32 = orc.contact_pairs(descs, verts, poses, pairs, 0.0)
    checked = 0
    for k in range(pairs.shape[0]):
        A = _world(verts[offs[2 * k]:offs[2 * k + 1]], poses[2 * k].astype(np.float64))
        B = _world(verts[offs[2 * k + 1]:offs[2 * k + 2]], poses[2 * k + 1].astype(np.float64))
        depth, axis = sat_depth(A, B)
        if depth < 1e-3:
            continue
        assert r64["status"][k] == 1 and r32["status"][k] == 1
        assert r64["iters"][k, 1] > 0  # EPA ran
        assert abs(-r64["out"][k, 12] - depth) <= 1e-7 * max(1.0, depth)
        assert abs(-r32["out"][k, 12] - depth) <= 1e-4 * max(1.0, depth)
        n1 = r64["out"][k, 6:9]
        assert np.dot(n1, axis) > 1 - 1e-6  # normal1 points from shape 1 towards shape 2
        checked += 1
    assert checked >= 20


def test_ball_ball_closed_form(orc):
    descs = np.concatenate([desc(SH_BALL, params=(0.75, 0, 0, 0)), desc(SH_BALL, params=(0.5, 0, 0, 0))])
    rng = np.random.default_rng(3)
    for _ in range(20):
        c2 = rng.uniform(-1.0, 1.0, 3)
        q1 = scenes.random_rotations(1, rng)[0]
        poses = np.array([list(q1) + [0, 0, 0], IDENT + list(c2)], np.float64)
        r = orc.contact_pairs(descs, np.zeros((0, 3)), poses, [[0, 1]], 0.05, f64=True)
        d = np.linalg.norm(c2)
        if d < 1.25 + 0.05:
            assert r["status"][0] == 1
            assert abs(r["out"][0, 12] - (d - 1.25)) < 1e-12
            assert np.allclose(r["out"][0, 6:9], c2 / d, atol=1e-12)  # world normal1
            assert np.allclose(r["out"][0, 0:3], c2 / d * 0.75, atol=1e-12)
            assert np.allclose(r["out"][0, 3:6], c2 - c2 / d * 0.5, atol=1e-12)
        else:
            assert r["status"][0] == 0


def test_f32_matches_f64_and_sign(orc):
    sc = scenes.polyhedra_scene(3000, seed=5)
    descs = np.zeros(3000, SHAPE_DESC_DTYPE)
    descs["type"] = SH_POLY
    descs["first_vertex"] = sc["offsets"][:-1]
    descs["num_vertices"] = sc["offsets"][1:] - sc["offsets"][:-1]
    aabbs = orc.polyhedra_aabbs(sc["vertices"], sc["offsets"], sc["poses"])
    pairs, _ = orc.OracleBvh(aabbs).self_pairs()
    r32 = orc.contact_pairs(descs, sc["vertices"], sc["poses"], pairs, 0.01)
    r64 = orc.contact_pairs(descs, sc["vertices"], sc["poses"], pairs, 0.01, f64=True)
    both = (r32["status"] == 1) & (r64["status"] == 1)
    assert (r32["status"] == r64["status"]).mean() > 0.99
    d32, d64 = r32["out"][both, 12], r64["out"][both, 12]
    assert np.quantile(np.abs(d32 - d64), 0.99) < 1e-3
    assert (np.sign(d32[np.abs(d64) > 1e-3]) == np.sign(d64[np.abs(d64) > 1e-3])).all()
