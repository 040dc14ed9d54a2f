"""Oracle pins for contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162)
on a fresh manifold.

  * hand-derived configurations: a small cuboid resting 0.1 deep on a large
    one's top face gives its four bottom corners (vertex ids of shape 2, the
    +z face id of shape 1) at dist -0.1 plus GJK's point; a capsule lying on a
    cuboid gives its two segment ends, moved out by the radius, at dist -0.05,
    plus GJK's point; two crossing capsules give one edge-edge point, two
    parallel ones the two ends of the overlap (clip_segment_segment);
  * face-face clipping against an independent computation: on random resting
    cuboid and cone-cap pairs, the clipped points (all but the last, GJK's) are
    exactly the vertices of the intersection of the two faces projected on the
    contact plane (polygon clipping in float64), to 1e-4;
  * every manifold's last point is GJK's and equals the penetration contact
    of the two feature maps;
  * separated pairs: status 0 with GJK's separating direction cached in n1;
    a Ball has no feature map (-1);
  * f32 against f64;
  * a box-shaped ConvexPolyhedron (faces from the hull) against a Cuboid of the
    same half extents in the same poses: the same statuses, normals and point
    sets (the clipping walks the faces from other vertices, so points agree to
    1e-4 but for crossings of nearly parallel edges), each vertex id naming the
    polyhedron vertex at the point and each face id a face whose plane holds it;
    a polyhedron without faces has no feature map;
  * the update from a previous frame (try_update_contacts_eps,
    contact_manifold.rs:662-699): a zero manifold gives the fresh result bit for
    bit; a 1e-4 rise keeps the points (status 2) with dist + 1e-4 and local_p1 the
    projection of the moved local_p2; a lateral move of 0.01, a 2 degree tilt and a
    point crossing from penetration to separation are recomputed (status 1); a
    separated pair's cached direction seeds GJK to the same separation.
"""
import numpy as np

from oracle import oracle as orc
from parry_amd.shape import SHAPE_DESC_DTYPE

S = np.float32(np.sqrt(0.5))
PF_VERTEX, PF_EDGE, PF_FACE = 1 << 30, 2 << 30, 3 << 30


def _descs(types, params):
    d = np.zeros(len(types), SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, p in enumerate(params):
        d["params"][i][: len(p)] = p
    return d


def _one(types, params, pose1, pose2, prediction=0.02, verts=None):
    d = _descs(types, params)
    v = np.zeros((0, 3), np.float32) if verts is None else verts
    return orc.manifolds_pfm(d, v, np.array([pose1, pose2], np.float32), np.array([[0, 1]], np.uint32), prediction)


def test_cuboid_resting_on_cuboid():
    r = _one([1, 1], [[1, 1, 1], [0.5, 0.5, 0.5]], [0, 0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0.2, 0.1, 1.4])
    assert r["status"][0] == 1 and r["count"][0] == 5
    np.testing.assert_allclose(r["n1"][0], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(r["n2"][0], [0, 0, -1], atol=1e-6)
    d = r["dist"][0, :4]
    np.testing.assert_allclose(d, -0.1, atol=1e-5)
    corners = {(round(float(x), 4), round(float(y), 4)) for x, y in r["p2"][0, :4, :2]}
    assert corners == {(0.5, 0.5), (0.5, -0.5), (-0.5, 0.5), (-0.5, -0.5)}
    np.testing.assert_allclose(r["p2"][0, :4, 2], -0.5, atol=1e-6)
    np.testing.assert_allclose(r["p1"][0, :4, 2], 1.0, atol=1e-5)
    assert (r["fid1"][0, :4] == PF_FACE | 15).all()  # +z face: iamax 2 + 3 * 1 + 10
    assert ((r["fid2"][0, :4] >> 30) == 1).all()  # vertices of shape 2
    assert r["fid1"][0, 4] == 0 and r["fid2"][0, 4] == 0  # GJK's point: unknown features
    np.testing.assert_allclose(r["dist"][0, 4], -0.1, atol=1e-5)


def test_capsule_lying_on_cuboid():
    # capsule axis (local y) turned onto x; segment 0.5 long each way, radius 0.2, 0.05 deep
    r = _one([1, 2], [[1, 1, 1], [0.5, 0.2]], [0, 0, 0, 1, 0, 0, 0], [0, 0, S, S, 0.1, 0.0, 1.15])
    assert r["status"][0] == 1 and r["count"][0] == 3
    np.testing.assert_allclose(r["n1"][0], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(r["dist"][0, :3], -0.05, atol=1e-5)
    # p2 in the capsule's frame: the segment ends (0, -+0.5, 0) moved by n2 * radius
    ends = sorted(float(y) for y in r["p2"][0, :2, 1])
    np.testing.assert_allclose(ends, [-0.5, 0.5], atol=1e-5)
    np.testing.assert_allclose(r["p1"][0, :2, 2], 1.0, atol=1e-5)  # on the cuboid's top face


def test_capsules_crossing_and_parallel():
    # crossing at right angles: one edge-edge point + GJK's
    r = _one([2, 2], [[0.5, 0.2], [0.5, 0.2]], [0, 0, S, S, 0, 0, 0], [0, 0, 0, 1, 0, 0, 0.35])
    assert r["status"][0] == 1 and r["count"][0] == 2
    assert (r["fid1"][0, 0] >> 30) == 2 and (r["fid2"][0, 0] >> 30) == 2
    np.testing.assert_allclose(r["dist"][0, :2], -0.05, atol=1e-5)
    # parallel, offset along the axis: the overlap's two ends + GJK's
    r = _one([2, 2], [[0.5, 0.2], [0.5, 0.2]], [0, 0, S, S, 0, 0, 0], [0, 0, S, S, 0.3, 0, 0.35])
    assert r["status"][0] == 1 and r["count"][0] == 3
    # the overlap is world x in [-0.2, 0.5]: local y (world -x) of the first capsule in [-0.5, 0.2]
    ys = sorted(float(y) for y in r["p1"][0, :2, 1])
    np.testing.assert_allclose(ys, [-0.5, 0.2], atol=1e-5)


def test_separated_and_unsupported():
    r = _one([1, 1], [[1, 1, 1], [0.5, 0.5, 0.5]], [0, 0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0, 3.0])
    assert r["status"][0] == 0 and r["count"][0] == 0
    # n1 caches GJK's separating direction: along it the boxes are more than the prediction apart
    n1 = r["n1"][0].astype(np.float64)
    assert abs(np.linalg.norm(n1) - 1) < 1e-6
    c1 = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    c2 = c1 * 0.5 + [0, 0, 3.0]
    assert (c2 @ n1).min() - (c1 @ n1).max() > 0.02
    r = _one([0, 1], [[0.5], [1, 1, 1]], [0, 0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0, 1.2])
    assert r["status"][0] == -1


def _rotate(q, v):
    q = np.asarray(q, np.float64)
    u, w = q[:3], q[3]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def _clip(poly, clipper):
    """Sutherland-Hodgman: `poly` clipped by the convex polygon `clipper` (both CCW or both CW)."""
    def side(a, b, p):
        return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0])

    orient = np.sign(sum(side(clipper[i], clipper[(i + 1) % len(clipper)], clipper[(i + 2) % len(clipper)])
                         for i in range(len(clipper))))
    out = list(poly)
    for i in range(len(clipper)):
        a, b = clipper[i], clipper[(i + 1) % len(clipper)]
        inp, out = out, []
        for k in range(len(inp)):
            p, q = inp[k], inp[(k + 1) % len(inp)]
            sp, sq = side(a, b, p) * orient, side(a, b, q) * orient
            if sp >= 0:
                out.append(p)
            if sp * sq < 0:
                t = sp / (sp - sq)
                out.append(p + (q - p) * t)
        if not out:
            break
    return out


def test_face_clipping_against_polygon_intersection():
    rng = np.random.default_rng(4)
    checked = 0
    for trial in range(150):
        he1 = rng.uniform(0.5, 1.5, 3)
        he2 = rng.uniform(0.3, 1.0, 3)
        ang = rng.uniform(-np.pi, np.pi)
        q2 = [0, 0, np.sin(ang / 2), np.cos(ang / 2)]  # yaw only: faces stay parallel
        pos2 = [rng.uniform(-1, 1), rng.uniform(-1, 1), he1[2] + he2[2] - rng.uniform(0.01, 0.1)]
        r = _one([1, 1], [list(he1), list(he2)], [0, 0, 0, 1, 0, 0, 0], q2 + pos2)
        if r["status"][0] != 1:
            continue
        n = int(r["count"][0])
        got = r["p1"][0, : n - 1, :2].astype(np.float64)  # the contact points on face 1, in its plane
        face1 = np.array([[he1[0], he1[1]], [he1[0], -he1[1]], [-he1[0], -he1[1]], [-he1[0], he1[1]]], np.float64)
        c2 = [_rotate(q2, np.array([sx * he2[0], sy * he2[1], -he2[2]])) + pos2 for sx, sy in
              ((1, 1), (1, -1), (-1, -1), (-1, 1))]
        face2 = np.array([[p[0], p[1]] for p in c2])
        inter = _clip(list(face1), list(face2))
        ref = np.array(inter)
        # every intersection vertex is a contact point and vice versa
        d1 = np.min(np.linalg.norm(got[:, None, :] - ref[None, :, :], axis=2), axis=1)
        d2 = np.min(np.linalg.norm(ref[:, None, :] - got[None, :, :], axis=2), axis=1)
        assert d1.max() < 1e-4 and d2.max() < 1e-4, (trial, got, ref)
        checked += 1
    assert checked > 100


def test_gjk_point_is_the_feature_maps_contact():
    rng = np.random.default_rng(8)
    n = 400
    types = rng.choice([1, 2, 3], size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, t in enumerate(types):
        d["params"][i][:3] = rng.uniform(0.3, 1.0, 3) if t == 1 else list(rng.uniform(0.3, 1.0, 2)) + [0]
    from parry_amd import scenes

    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[1::2, 4:] = rng.normal(0, 0.6, (n, 3))
    poses = poses.astype(np.float32)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    r = orc.manifolds_pfm(d, np.zeros((0, 3), np.float32), poses, pairs, 0.02)
    ok = r["status"] == 1
    assert ok.sum() > 150
    # without capsules (no border), GJK's point is query::contact's in local frames
    nocap = ok & (types[pairs[:, 0]] != 2) & (types[pairs[:, 1]] != 2)
    c = orc.contact_pairs(d, np.zeros((0, 3), np.float32), poses, pairs, 0.02)
    idx = np.nonzero(nocap)[0]
    for k in idx:
        last = int(r["count"][k]) - 1
        i = pairs[k, 0]
        p1w = _rotate(poses[i, :4], r["p1"][k, last].astype(np.float64)) + poses[i, 4:]
        np.testing.assert_allclose(p1w, c["out"][k, 0:3], atol=2e-5)
        np.testing.assert_allclose(r["dist"][k, last], c["out"][k, 12], atol=2e-5)
    assert len(idx) > 30


def test_f32_agrees_with_f64():
    rng = np.random.default_rng(2)
    n = 500
    types = rng.choice([1, 2, 3], size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, t in enumerate(types):
        d["params"][i][:3] = rng.uniform(0.3, 1.0, 3) if t == 1 else list(rng.uniform(0.3, 1.0, 2)) + [0]
    from parry_amd import scenes

    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[1::2, 4:] = rng.normal(0, 0.7, (n, 3))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    a = orc.manifolds_pfm(d, np.zeros((0, 3)), poses, pairs, 0.02)
    b = orc.manifolds_pfm(d, np.zeros((0, 3)), poses, pairs, 0.02, f64=True)
    assert (a["status"] == b["status"]).mean() > 0.98
    same = (a["status"] == 1) & (b["status"] == 1) & (a["count"] == b["count"])
    assert same.mean() > 0.6
    dd = np.abs(a["dist"][same] - b["dist"][same])
    assert np.quantile(dd, 0.99) < 1e-3


def _box_polyhedra(he):
    signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    verts = (he[:, None, :] * signs[None]).reshape(-1, 3).astype(np.float32)
    d = np.zeros(len(he), SHAPE_DESC_DTYPE)
    d["type"] = 4
    d["first_vertex"] = 8 * np.arange(len(he))
    d["num_vertices"] = 8
    return d, verts


def test_box_polyhedron_matches_cuboid():
    from parry_amd import scenes

    rng = np.random.default_rng(5)
    n = 2000
    dp, verts = _box_polyhedra(rng.uniform(0.3, 1.0, (2 * n, 3)))
    dc = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    dc["type"] = 1
    dc["params"][:, :3] = verts.reshape(-1, 8, 3)[:, 7]
    faces = scenes.polyhedron_faces(dp, verts)
    assert (faces["shape_faces"][:, 1] == 6).all() and (faces["faces"]["count"] == 4).all()
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[1::2, 4:] = rng.normal(0, 0.8, (n, 3))
    poses = poses.astype(np.float32)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    a = orc.manifolds_pfm(dp, verts, poses, pairs, 0.02, faces=faces)
    b = orc.manifolds_pfm(dc, np.zeros((0, 3), np.float32), poses, pairs, 0.02)
    assert np.array_equal(a["status"], b["status"]) and (a["status"] == 1).sum() > 1000
    ok = a["status"] == 1
    assert np.array_equal(a["count"][ok], b["count"][ok])
    np.testing.assert_allclose(a["n1"][ok], b["n1"][ok], atol=1e-5)
    close, checked_ids = 0, 0
    fn = faces["faces"]["normal"].astype(np.float64)
    for k in np.nonzero(ok)[0]:
        c = int(a["count"][k])
        pa = np.concatenate([a["p1"][k, :c], a["p2"][k, :c], a["dist"][k, :c, None]], 1)
        pb = np.concatenate([b["p1"][k, :c], b["p2"][k, :c], b["dist"][k, :c, None]], 1)
        e = np.abs(pa[:, None, :] - pb[None, :, :]).max(2).min(1).max()
        assert e < 1e-2, f"pair {k}: points differ by {e}"
        close += e <= 1e-4
        for side, (fids, pts) in enumerate(((a["fid1"][k, :c], a["p1"][k, :c]), (a["fid2"][k, :c], a["p2"][k, :c]))):
            s_idx = int(pairs[k, side])
            for fid, p in zip(fids, pts):
                kind, idx = int(fid) >> 30, int(fid) & ((1 << 30) - 1)
                if kind == 1:  # vertex: the point is that vertex
                    np.testing.assert_allclose(p, verts[8 * s_idx + idx], atol=1e-5)
                    checked_ids += 1
                elif kind == 3:  # face: the point lies in that face's plane
                    f = faces["shape_faces"][s_idx, 0] + idx
                    v0 = verts[8 * s_idx + faces["face_vertices"][faces["faces"]["first"][f]]]
                    assert abs(np.dot(fn[f], p.astype(np.float64) - v0)) < 1e-5
                    checked_ids += 1
    assert close >= 0.995 * ok.sum() and checked_ids > 2000
    # no faces: no feature map
    r = orc.manifolds_pfm(dp[:2], verts[:16], poses[:2], np.array([[0, 1]], np.uint32), 0.02)
    assert r["status"][0] == -1


def _frame(d, poses, prev, verts=None):
    v = np.zeros((0, 3), np.float32) if verts is None else verts
    return orc.manifolds_pfm(d, v, np.asarray(poses, np.float32), np.array([[0, 1]], np.uint32), 0.02, previous=prev)


def test_manifold_update_keeps_or_recomputes():
    # try_update_contacts_eps (contact_manifold.rs:662-699) on the resting cuboid of
    # test_cuboid_resting_on_cuboid, then seeded recomputation when it refuses
    d = _descs([1, 1], [[1, 1, 1], [0.5, 0.5, 0.5]])
    base = [0, 0, 0, 1, 0.2, 0.1, 1.4]
    r0 = _frame(d, [[0, 0, 0, 1, 0, 0, 0], base], None)
    assert r0["status"][0] == 1 and r0["count"][0] == 5
    # a fresh manifold through the update path is the fresh result
    zero = {k: np.zeros_like(v) for k, v in r0.items()}
    rz = _frame(d, [[0, 0, 0, 1, 0, 0, 0], base], zero)
    for k in ("status", "count", "n1", "n2", "p1", "p2", "dist", "fid1", "fid2"):
        assert np.array_equal(rz[k], r0[k]), k
    # shape 2 rises by 1e-4: every point moves by less than 1e-3, the normal is
    # unchanged: kept, dist + 1e-4, local_p1 the projection of the moved local_p2
    up = list(base)
    up[6] += 1e-4
    r1 = _frame(d, [[0, 0, 0, 1, 0, 0, 0], up], r0)
    assert r1["status"][0] == 2 and r1["count"][0] == 5
    for k in ("n1", "n2", "p2", "fid1", "fid2"):
        assert np.array_equal(r1[k], r0[k]), k
    c = 5
    dz = np.float32(up[6]) - np.float32(base[6])
    np.testing.assert_allclose(r1["dist"][0, :c], r0["dist"][0, :c] + dz, atol=1e-6)
    lp2 = r0["p2"][0, :c].astype(np.float64) + np.array([0.2, 0.1, float(np.float32(up[6]))])
    np.testing.assert_allclose(r1["p1"][0, :c], lp2 - np.outer(r1["dist"][0, :c], [0, 0, 1]), atol=1e-6)
    # a lateral move of 0.01 moves each point by 1e-4 squared > 1e-6: recomputed (status 1),
    # as a fresh manifold would be; GJK's own point (the last) is any point of the
    # overlap and depends on the seed, the four clipped corners do not
    side = list(base)
    side[4] += 0.01
    r2 = _frame(d, [[0, 0, 0, 1, 0, 0, 0], side], r0)
    fresh = _frame(d, [[0, 0, 0, 1, 0, 0, 0], side], None)
    assert r2["status"][0] == 1 and r2["count"][0] == fresh["count"][0] == 5
    np.testing.assert_allclose(r2["p1"][0, :4], fresh["p1"][0, :4], atol=1e-5)
    np.testing.assert_allclose(r2["dist"][0, :c], fresh["dist"][0, :c], atol=1e-5)
    # a 2 degree turn of shape 2 about z keeps the normal (rotation about it) but
    # one about x tilts it beyond cos(1 degree): recomputed
    a = np.radians(2.0) / 2
    tilt = [np.sin(a), 0, 0, np.cos(a), 0.2, 0.1, 1.4]
    r3 = _frame(d, [[0, 0, 0, 1, 0, 0, 0], tilt], r0)
    assert r3["status"][0] == 1
    # penetration to separation by 2e-5 on a point at -1e-5: recomputed even
    # though the points barely move
    near = list(base)
    near[6] = 1.5 - 1e-5
    rn = _frame(d, [[0, 0, 0, 1, 0, 0, 0], near], None)
    assert rn["status"][0] == 1 and (rn["dist"][0, : rn["count"][0]] < 0).all()
    away = list(near)
    away[6] += 2e-5
    ra = _frame(d, [[0, 0, 0, 1, 0, 0, 0], away], rn)
    assert ra["status"][0] == 1
    rb = _frame(d, [[0, 0, 0, 1, 0, 0, 0], near], rn)  # no motion: kept
    assert rb["status"][0] == 2
    # separated: the cached direction seeds GJK, which finds the same separation
    far = [0, 0, 0, 1, 0.2, 0.1, 1.6]
    rs = _frame(d, [[0, 0, 0, 1, 0, 0, 0], far], None)
    assert rs["status"][0] == 0
    rs2 = _frame(d, [[0, 0, 0, 1, 0, 0, 0], far], rs)
    assert rs2["status"][0] == 0
    np.testing.assert_allclose(rs2["n1"][0], [0, 0, 1], atol=1e-5)  # GJK's separating direction
