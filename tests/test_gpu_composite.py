"""GPU parity of contact between a TriMesh and convex shapes (C ABI).

Bar, against the f32 oracle (contact_composite_shape_shape.rs:14-59 through
DefaultQueryDispatcher::contact):
  * has_contact identical for every shape, and the index of the triangle in
    contact identical (PQ_NO_HIT for None);
  * point1, point2, normal1, normal2 and dist within 1e-5 relative
    (|a-b| <= 1e-5 * max(1, |b|));
  * the number of (triangle, shape) candidates identical: the Bvh walk selects
    the same leaves;
  * with a candidate capacity below what the call needs, every shape with a
    candidate past the capacity reads PQ_CONTACT_TRUNCATED and every other
    shape is unchanged;
plus the flat-plane closed forms, and Ball / Triangle shapes in the batched
pair contact.
"""
import numpy as np
import pytest

from parry_amd import (
    PQ_CONTACT_TRUNCATED,
    SHAPE_DESC_DTYPE,
    MeshContact,
    ShapeSet,
    TriMesh,
    contact_pairs,
    scenes,
)

pytestmark = pytest.mark.gpu
TOL = 1e-5
NO_HIT = 0xFFFFFFFF
IDENT = np.array([0, 0, 0, 1, 0, 0, 0], np.float32)


def _run(mc, mesh_pose, poses, prediction, max_candidates):
    import torch

    mp = torch.from_numpy(np.ascontiguousarray(mesh_pose, np.float32)).cuda()
    p = torch.from_numpy(np.ascontiguousarray(poses, np.float32)).cuda()
    out, tri, needed = mc.contact(mp, p, prediction, max_candidates)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[:, :13], o.view(np.uint32)[:, 13].copy(), tri.cpu().numpy().view(np.uint32), int(needed.item())


def _compare(label, got, ref):
    vals, has, tri, needed = got
    st = ref["status"].astype(np.uint32)
    assert np.array_equal(has, st), f"{label}: {np.count_nonzero(has != st)} has_contact mismatches"
    c = st == 1
    assert np.array_equal(tri[c], ref["triangle"][c]), (
        f"{label}: {np.count_nonzero(tri[c] != ref['triangle'][c])} triangle indices differ")
    assert (tri[~c] == NO_HIT).all()
    err = np.abs(vals[c] - ref["out"][c]) / np.maximum(1.0, np.abs(ref["out"][c]))
    exact = np.all(vals[c] == ref["out"][c], axis=1)
    print(f"{label}: {c.sum()} of {len(st)} shapes in contact, {exact.sum()} bit-identical, "
          f"max rel err {err.max(initial=0):.3e}, {needed} candidates", flush=True)
    assert err.max(initial=0) <= TOL, f"{label}: {(err > TOL).any(axis=1).sum()} contacts beyond {TOL}"
    assert needed == ref["candidates"], f"{label}: {needed} candidates, oracle {ref['candidates']}"
    return c


def _pose(axis, angle, t):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.concatenate([a * np.sin(angle / 2), [np.cos(angle / 2)], t]).astype(np.float32)


def test_icosphere_shapes_match_oracle(orc):
    v, f = scenes.noisy_icosphere(5, seed=1, radius=2.0)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    descs, verts, poses = scenes.shapes_around_mesh(20000, seed=13)
    mc = MeshContact(mesh, ShapeSet(descs, verts))
    for label, mp in (("identity", IDENT), ("posed", _pose([1, 2, -0.5], 0.7, [0.3, -0.2, 0.1]))):
        for pred in (0.0, 0.05):
            print(f"case: mesh contact, icosphere s=5 r=2, 20000 mixed shapes, mesh pose {label}, "
                  f"prediction {pred}", flush=True)
            ref = om.contact_shapes(mp, descs, verts, poses, pred)
            c = _compare(f"{label} pred={pred}", _run(mc, mp, poses, pred, ref["candidates"]), ref)
            assert c.sum() > 2000 and (~c).sum() > 2000


def test_non_default_epa_tunings(orc):
    """The mesh contact's EPA pairs (a triangle against a shape) under first-tier
    settings other than the default: wave batches with 8 lanes per pair and the
    64-face polytope, and the per-group refill with 32 lanes per pair."""
    from parry_amd import _native

    v, f = scenes.noisy_icosphere(4, seed=2, radius=2.0)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    descs, verts, poses = scenes.shapes_around_mesh(8000, seed=14)
    mc = MeshContact(mesh, ShapeSet(descs, verts))
    ref = om.contact_shapes(IDENT, descs, verts, poses, 0.02)
    saved = _native.get_contact_tuning()
    try:
        for g, poly, sched in ((8, 2, 1), (32, 0, 2)):
            _native.set_contact_tuning(g, poly, sched)
            print(f"case: mesh contact, icosphere s=4, 8000 shapes, EPA G={g} polytope={poly} schedule={sched}",
                  flush=True)
            c = _compare(f"G={g} polytope={poly} schedule={sched}", _run(mc, IDENT, poses, 0.02, ref["candidates"]),
                         ref)
            assert c.sum() > 500
    finally:
        _native.set_contact_tuning(*saved)


def test_truncation(orc):
    v, f = scenes.noisy_icosphere(4, seed=2, radius=2.0)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    descs, verts, poses = scenes.shapes_around_mesh(400, seed=21)
    mc = MeshContact(mesh, ShapeSet(descs, verts))
    print("case: mesh contact truncation, icosphere s=4, 400 shapes", flush=True)
    full = om.contact_shapes(IDENT, descs, verts, poses, 0.02)
    counts = np.array([om.contact_shapes(IDENT, descs[q:q + 1], verts, poses[q:q + 1], 0.02)["candidates"]
                       for q in range(len(descs))], np.int64)
    assert counts.sum() == full["candidates"]
    ends = np.cumsum(counts)
    for cap in (0, 1, full["candidates"] // 3, full["candidates"] - 1, full["candidates"]):
        vals, has, tri, needed = _run(mc, IDENT, poses, 0.02, cap)
        assert needed == full["candidates"]
        trunc = (ends > cap) & (counts > 0)
        assert (has[trunc] == PQ_CONTACT_TRUNCATED).all(), f"cap {cap}"
        keep = ~trunc
        assert np.array_equal(has[keep], full["status"][keep].astype(np.uint32)), f"cap {cap}"
        c = keep & (full["status"] == 1)
        assert np.array_equal(tri[c], full["triangle"][c])
        assert np.array_equal(vals[c], _run(mc, IDENT, poses, 0.02, full["candidates"])[0][c])
        print(f"  capacity {cap}: {trunc.sum()} truncated", flush=True)


def _descs(types, params):
    d = np.zeros(len(types), SHAPE_DESC_DTYPE)
    d["type"] = types
    for i, p in enumerate(params):
        d["params"][i][: len(p)] = p
    return d


def test_plane_closed_forms(orc):
    pv = np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32)
    pf = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    s = np.float32(np.sqrt(0.5))
    d = _descs([0, 1, 2, 3], [[0.5], [0.2, 0.2, 0.2], [0.3, 0.1], [0.5, 0.4]])
    poses = np.array([[0, 0, 0, 1, 0.3, 0.2, 0.4], [0, 0, 0, 1, 0.1, -0.3, 0.15],
                      [s, 0, 0, s, -1.0, 1.0, 0.35], [s, 0, 0, s, 1.5, 1.5, 0.45]], np.float32)
    nov = np.zeros((0, 3), np.float32)
    mc = MeshContact(TriMesh(pv, pf), ShapeSet(d, nov))
    om = orc.OracleTriMesh(pv, pf)
    print("case: mesh contact, flat plane, Ball/Cuboid/Capsule/Cone", flush=True)
    got = _run(mc, IDENT, poses, 0.01, 64)
    _compare("plane", got, om.contact_shapes(IDENT, d, nov, poses, 0.01))
    vals = got[0]
    np.testing.assert_allclose(vals[:, 12], [-0.1, -0.05, -0.05, -0.05], atol=2e-5)
    np.testing.assert_allclose(vals[:, 6:9], np.tile([0, 0, 1], (4, 1)), atol=1e-5)
    # the ball rests across the shared diagonal edge of the two triangles: the
    # first triangle in leaf order wins the tie
    edge = poses.copy()
    edge[0, 4:7] = [0.5, 0.5, 0.4]
    _compare("plane edge", _run(mc, IDENT, edge, 0.01, 64), om.contact_shapes(IDENT, d, nov, edge, 0.01))
    far = poses + np.array([0, 0, 0, 0, 0, 0, 1.0], np.float32)
    vals, has, tri, needed = _run(mc, IDENT, far, 0.01, 64)
    assert (has == 0).all() and (tri == NO_HIT).all() and needed == 0


def test_triangle_shapes_in_pair_contact(orc):
    import torch

    rng = np.random.default_rng(5)
    n = 3000
    types = rng.choice([0, 1, 2, 3, 5], size=n)
    d = np.zeros(n, SHAPE_DESC_DTYPE)
    d["type"] = types
    tri_verts = []
    for i, t in enumerate(types):
        if t == 0:
            d["params"][i][0] = rng.uniform(0.2, 0.6)
        elif t == 1:
            d["params"][i][:3] = rng.uniform(0.2, 0.6, 3)
        elif t in (2, 3):
            d["params"][i][:2] = rng.uniform(0.2, 0.6, 2)
        else:
            d["first_vertex"][i] = 3 * len(tri_verts)
            d["num_vertices"][i] = 3
            tri_verts.append(rng.normal(0.0, 0.6, (3, 3)))
    verts = np.concatenate(tri_verts, 0).astype(np.float32)
    q = scenes.random_rotations(n, rng)
    t = rng.uniform(0.0, 9.0, (n, 3))
    poses = np.concatenate([q, t], 1).astype(np.float32)
    # pairs of nearby shapes, every one with a triangle on one side or both
    ctr = poses[:, 4:]
    dd = np.linalg.norm(ctr[:, None, :] - ctr[None, :, :], axis=2)
    i, j = np.nonzero((dd < 1.4) & (np.arange(n)[:, None] != np.arange(n)[None, :]))
    keep = (types[i] == 5) | (types[j] == 5)
    pairs = np.stack([i[keep], j[keep]], 1).astype(np.uint32)
    print(f"case: pair contact with Triangle shapes, {pairs.shape[0]} pairs", flush=True)
    shapes = ShapeSet(d, verts)
    out = contact_pairs(shapes, torch.from_numpy(poses).cuda(), torch.from_numpy(pairs.view(np.int32)).cuda(), 0.02)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    ref = orc.contact_pairs(d, verts, poses, pairs, 0.02)
    has = o.view(np.uint32)[:, 13]
    assert (ref["status"] >= 0).all()
    assert np.array_equal(has.astype(np.int64), ref["status"].astype(np.int64))
    c = ref["status"] == 1
    err = np.abs(o[c, :13] - ref["out"][c]) / np.maximum(1.0, np.abs(ref["out"][c]))
    kinds = {(a, b) for a, b in zip(types[pairs[c, 0]], types[pairs[c, 1]])}
    print(f"  {c.sum()} contacts over {len(kinds)} type pairs, max rel err {err.max(initial=0):.3e}", flush=True)
    assert err.max(initial=0) <= TOL
    assert (0, 5) in kinds and (5, 0) in kinds and (5, 5) in kinds and (5, 1) in kinds
