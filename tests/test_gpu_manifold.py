"""GPU parity of contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162)
on fresh manifolds against the oracle, through the C ABI.

Bar: status, number of points and every feature id identical to the f32
oracle; local_n1, local_n2 and every point's local_p1, local_p2 and dist within
1e-5 relative (|a-b| <= 1e-5 * max(1, |b|)); the number of bit-identical
manifolds is printed.  Cases: random mixes of Cuboid, Capsule, Cone and
Triangle in contact (face-face, edge-edge and mixed clipping, EPA for the
penetrating ones), resting stacks, separated pairs (the cached direction), and
pairs with a Ball or a ConvexPolyhedron without faces (unsupported),
ConvexPolyhedra given their faces (random hulls, boxes) against each other and
the other feature maps, and four frames of the update from the previous frame's
manifolds (pq_contact_manifolds_pfm_update) with shapes at rest, barely moving
and moving: each frame, the oracle starts from the GPU's previous manifolds.
"""
import numpy as np
import pytest

from parry_amd import CONTACT_MANIFOLD_DTYPE, SHAPE_DESC_DTYPE, ShapeSet, contact_manifolds_pfm, scenes

pytestmark = pytest.mark.gpu
TOL = 1e-5
UNSUP = 0xFFFFFFFF


def _scene(n, seed, spread, types=(1, 2, 3, 5)):
    rng = np.random.default_rng(seed)
    t = rng.choice(types, size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = t
    tv = []
    for i, ty in enumerate(t):
        if ty == 1:
            d["params"][i][:3] = rng.uniform(0.3, 1.0, 3)
        elif ty in (2, 3):
            d["params"][i][:2] = rng.uniform(0.3, 1.0, 2)
        elif ty == 5:
            d["first_vertex"][i] = 3 * len(tv)
            d["num_vertices"][i] = 3
            tv.append(rng.normal(0, 0.7, (3, 3)))
        elif ty == 0:
            d["params"][i][0] = rng.uniform(0.3, 1.0)
    verts = np.concatenate(tv, 0).astype(np.float32) if tv else np.zeros((0, 3), np.float32)
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[0::2, 4:] = rng.uniform(-40, 40, (n, 3))
    poses[1::2, 4:] = poses[0::2, 4:] + rng.normal(0, spread, (n, 3))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    return d, verts, poses.astype(np.float32), pairs


def _run(d, verts, poses, pairs, prediction, faces=None):
    import torch

    shapes = ShapeSet(d, verts)
    if faces is not None:
        shapes.set_polyhedron_faces(faces)
    out = contact_manifolds_pfm(shapes, torch.from_numpy(poses).cuda(),
                                torch.from_numpy(pairs.view(np.int32)).cuda(), prediction)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(CONTACT_MANIFOLD_DTYPE).reshape(-1)


def _compare(label, got, ref):
    st = np.where(ref["status"] < 0, UNSUP, ref["status"]).astype(np.uint32)
    assert np.array_equal(got["status"], st), f"{label}: {np.count_nonzero(got['status'] != st)} status mismatches"
    assert np.array_equal(got["num_points"], ref["count"]), f"{label}: point counts differ"
    ok = st != UNSUP
    for name, a, b in (("n1", got["local_n1"][ok], ref["n1"][ok]), ("n2", got["local_n2"][ok], ref["n2"][ok])):
        err = np.abs(a - b) / np.maximum(1.0, np.abs(b))
        assert err.max(initial=0) <= TOL, f"{label}: {name} error {err.max()}"
    exact = 0
    worst = 0.0
    for k in np.nonzero((st == 1) | (st == 2))[0]:
        c = int(ref["count"][k])
        p = got["points"][k, :c]
        assert np.array_equal(p["fid1"], ref["fid1"][k, :c]) and np.array_equal(p["fid2"], ref["fid2"][k, :c]), (
            f"{label}: pair {k} feature ids differ")
        vals = np.concatenate([p["local_p1"], p["local_p2"], p["dist"][:, None]], 1)
        rv = np.concatenate([ref["p1"][k, :c], ref["p2"][k, :c], ref["dist"][k, :c, None]], 1)
        err = np.abs(vals - rv) / np.maximum(1.0, np.abs(rv))
        worst = max(worst, float(err.max(initial=0)))
        exact += int(np.array_equal(vals, rv))
    npts = int(ref["count"][(st == 1) | (st == 2)].sum())
    print(f"{label}: {int((st == 1).sum())} manifolds + {int((st == 2).sum())} kept ({npts} points), {exact} bit-identical, "
          f"{int((st == 0).sum())} separated, {int((st == UNSUP).sum())} unsupported, max rel err {worst:.3e}",
          flush=True)
    assert worst <= TOL
    return st


def test_random_pairs_match_oracle(orc):
    for seed, spread in ((3, 0.6), (4, 1.0)):
        d, verts, poses, pairs = _scene(40000, seed, spread)
        print(f"case: manifolds, 40000 pairs seed {seed} spread {spread}", flush=True)
        ref = orc.manifolds_pfm(d, verts, poses, pairs, 0.02)
        st = _compare(f"seed {seed}", _run(d, verts, poses, pairs, 0.02), ref)
        assert (st == 1).sum() > 5000 and (st == 0).sum() > 1000


def test_resting_stacks_match_oracle(orc):
    # cuboids and cone caps resting face down on cuboids: face-face clipping with many points
    rng = np.random.default_rng(6)
    n = 20000
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"][0::2] = 1
    d["type"][1::2] = rng.choice([1, 3], size=n)
    d["params"][0::2, :3] = rng.uniform(0.5, 1.5, (n, 3))
    d["params"][1::2, :3] = rng.uniform(0.2, 0.8, (n, 3))
    ang = rng.uniform(-np.pi, np.pi, n)
    poses = np.zeros((2 * n, 7), np.float32)
    poses[0::2, 3] = 1.0
    poses[0::2, 4:] = rng.uniform(-50, 50, (n, 3))
    cone = d["type"][1::2] == 3
    # cuboids: yaw about z; cones: their axis (y) turned onto +z, cap (y = -h) facing down
    poses[1::2, 2] = np.where(cone, 0.0, np.sin(ang / 2))
    poses[1::2, 3] = np.where(cone, np.sqrt(0.5), np.cos(ang / 2))
    poses[1::2, 0] = np.where(cone, np.sqrt(0.5), 0.0)
    h2 = np.where(cone, d["params"][1::2, 0], d["params"][1::2, 2])
    poses[1::2, 4:6] = poses[0::2, 4:6] + rng.uniform(-0.8, 0.8, (n, 2))
    poses[1::2, 6] = poses[0::2, 6] + d["params"][0::2, 2] + h2 - rng.uniform(0.0, 0.05, n)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    print("case: manifolds, 20000 resting cuboids and cones", flush=True)
    ref = orc.manifolds_pfm(d, np.zeros((0, 3), np.float32), poses, pairs, 0.02)
    st = _compare("resting", _run(d, np.zeros((0, 3), np.float32), poses, pairs, 0.02), ref)
    assert (ref["count"][st == 1] >= 4).mean() > 0.5


def test_unsupported_shapes(orc):
    d, verts, poses, pairs = _scene(5000, 9, 0.8, types=(0, 1, 2, 3))
    print("case: manifolds with balls (no feature map)", flush=True)
    ref = orc.manifolds_pfm(d, verts, poses, pairs, 0.02)
    st = _compare("with balls", _run(d, verts, poses, pairs, 0.02), ref)
    assert (st == UNSUP).sum() > 1000


def _polyhedra_scene(n, seed, spread):
    """Pairs over Cuboid, Capsule, Cone, Triangle and ConvexPolyhedron (half random
    hulls of 8..40 points, half boxes), polyhedra in at least one slot of most pairs."""
    rng = np.random.default_rng(seed)
    t = rng.choice([4, 4, 4, 1, 2, 3, 5], size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = t
    pts = []
    nv = 0
    for i, ty in enumerate(t):
        if ty == 1:
            d["params"][i][:3] = rng.uniform(0.3, 1.0, 3)
        elif ty in (2, 3):
            d["params"][i][:2] = rng.uniform(0.3, 1.0, 2)
        else:
            if ty == 5:
                v = rng.normal(0, 0.7, (3, 3))
            elif rng.uniform() < 0.5:
                v = scenes.convex_hull_vertices(scenes.convex_point_cloud(rng, int(rng.integers(8, 41))))
            else:
                he = rng.uniform(0.3, 1.0, 3)
                v = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * he
            d["first_vertex"][i] = nv
            d["num_vertices"][i] = len(v)
            pts.append(v)
            nv += len(v)
    verts = np.concatenate(pts, 0).astype(np.float32)
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[0::2, 4:] = rng.uniform(-40, 40, (n, 3))
    poses[1::2, 4:] = poses[0::2, 4:] + rng.normal(0, spread, (n, 3))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    return d, verts, poses.astype(np.float32), pairs


def test_polyhedra_with_faces_match_oracle(orc):
    d, verts, poses, pairs = _polyhedra_scene(6000, 11, 0.8)
    faces = scenes.polyhedron_faces(d, verts)
    print(f"case: manifolds, 6000 pairs with ConvexPolyhedra ({len(faces['faces'])} faces)", flush=True)
    ref = orc.manifolds_pfm(d, verts, poses, pairs, 0.02, faces=faces)
    st = _compare("polyhedra", _run(d, verts, poses, pairs, 0.02, faces=faces), ref)
    poly = (d["type"][pairs[:, 0]] == 4) | (d["type"][pairs[:, 1]] == 4)
    assert (st[poly] == 1).sum() > 1500 and (st == UNSUP).sum() == 0
    print("case: the same pairs, polyhedra without faces", flush=True)
    st = _compare("no faces", _run(d, verts, poses, pairs, 0.02), orc.manifolds_pfm(d, verts, poses, pairs, 0.02))
    assert (st[poly] == UNSUP).all()


def _as_oracle(rec):
    """GPU manifold records -> the oracle's manifold dict (its `previous` input)."""
    st = rec["status"].astype(np.int64)
    st[st == UNSUP] = -1
    pts = rec["points"]
    return {"status": st.astype(np.int32), "n1": rec["local_n1"].copy(), "n2": rec["local_n2"].copy(),
            "count": rec["num_points"].copy(), "p1": pts["local_p1"].copy(), "p2": pts["local_p2"].copy(),
            "dist": pts["dist"].copy(), "fid1": pts["fid1"].copy(), "fid2": pts["fid2"].copy()}


def _move(poses, rng):
    """A third of the shapes rest, a third move by ~1e-4, a third by ~0.02 and turn a little."""
    m = poses.shape[0]
    kind = rng.integers(0, 3, m)
    step = np.where(kind == 1, 1e-4, np.where(kind == 2, 0.02, 0.0))[:, None]
    poses[:, 4:] += (rng.normal(0, 1, (m, 3)) * step).astype(np.float32)
    turn = kind == 2
    dq = np.concatenate([rng.normal(0, 0.01, (int(turn.sum()), 3)), np.ones((int(turn.sum()), 1))], 1)
    q = poses[turn, :4].astype(np.float64)
    # q <- dq * q (x, y, z, w), renormalized
    x1, y1, z1, w1 = dq.T
    x2, y2, z2, w2 = q.T
    nq = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                   w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
    poses[turn, :4] = (nq / np.linalg.norm(nq, axis=1, keepdims=True)).astype(np.float32)


def test_update_from_previous_frames_matches_oracle(orc):
    import torch

    from parry_amd import contact_manifolds_pfm_update

    n = 30000
    d, verts, poses, pairs = _scene(n, 17, 0.6, types=(0, 1, 2, 3, 5))
    shapes = ShapeSet(d, verts)
    tp = torch.from_numpy(pairs.view(np.int32)).cuda()
    mf = torch.zeros((n, CONTACT_MANIFOLD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(5)
    prev = None
    kept_total = 0
    for frame in range(4):
        if frame:
            _move(poses, rng)
        print(f"case: manifold update, frame {frame}, {n} pairs", flush=True)
        ref = orc.manifolds_pfm(d, verts, poses, pairs, 0.02, previous=prev if prev is not None else
                                _as_oracle(np.zeros(n, CONTACT_MANIFOLD_DTYPE)))
        contact_manifolds_pfm_update(shapes, torch.from_numpy(poses).cuda(), tp, 0.02, mf)
        torch.cuda.synchronize()
        got = mf.cpu().numpy().view(CONTACT_MANIFOLD_DTYPE).reshape(-1)
        st = _compare(f"frame {frame}", got, ref)
        kept_total += int((st == 2).sum())
        if frame == 0:
            assert (st == 2).sum() == 0
        prev = _as_oracle(got)
    assert kept_total > 5000


def test_update_polyhedra_with_faces_matches_oracle(orc):
    """The update path with ConvexPolyhedra that have face tables (kept manifolds
    carry polyhedron feature ids across frames; separated pairs keep last
    frame's local_n2)."""
    import torch

    from parry_amd import contact_manifolds_pfm_update

    n = 6000
    d, verts, poses, pairs = _polyhedra_scene(n, 12, 0.8)
    faces = scenes.polyhedron_faces(d, verts)
    shapes = ShapeSet(d, verts)
    shapes.set_polyhedron_faces(faces)
    tp = torch.from_numpy(pairs.view(np.int32)).cuda()
    mf = torch.zeros((n, CONTACT_MANIFOLD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(8)
    prev = _as_oracle(np.zeros(n, CONTACT_MANIFOLD_DTYPE))
    kept_total = 0
    for frame in range(4):
        if frame:
            _move(poses, rng)
        print(f"case: manifold update with polyhedron faces, frame {frame}, {n} pairs", flush=True)
        ref = orc.manifolds_pfm(d, verts, poses, pairs, 0.02, faces=faces, previous=prev)
        contact_manifolds_pfm_update(shapes, torch.from_numpy(poses).cuda(), tp, 0.02, mf)
        torch.cuda.synchronize()
        got = mf.cpu().numpy().view(CONTACT_MANIFOLD_DTYPE).reshape(-1)
        st = _compare(f"polyhedra frame {frame}", got, ref)
        kept_total += int((st == 2).sum())
        prev = _as_oracle(got)
    assert kept_total > 1000
