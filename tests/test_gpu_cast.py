"""GPU parity of batched query::cast_shapes against the oracle (C ABI).

Bar: the status (None, Converged, PenetratingOrWithinTargetDist) identical to
the f32 oracle on every pair, the GJK ray-cast iteration count identical, and
toi, witness1, witness2, normal1 and normal2 within 1e-5 relative
(|a-b| <= 1e-5 * max(1, |b|)); the number of bit-identical records is printed.
Cases: mixed shapes (Ball, Cuboid, Capsule, Cone, ConvexPolyhedron, Triangle)
in approach, under every option (target distance, stop_at_penetration,
compute_impact_geometry_on_penetration, a finite max_time_of_impact), pairs that
start in penetration (the impact geometry runs through GJK and EPA), and the
reference's own shape-cast tests.
"""
import numpy as np
import pytest

from parry_amd import SHAPE_DESC_DTYPE, ShapeSet, cast_shapes_pairs, scenes

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _run(shapes, poses, vels, pairs, **opts):
    import torch

    p = torch.from_numpy(poses).cuda()
    v = torch.from_numpy(vels).cuda()
    pr = torch.from_numpy(pairs.view(np.int32)).cuda()
    out, iters = cast_shapes_pairs(shapes, p, v, pr, with_iters=True, **opts)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[:, :13], o.view(np.uint32)[:, 13].copy(), iters.cpu().numpy().view(np.uint32)


def _compare(label, got, ref):
    vals, st, iters = got
    assert np.array_equal(st, ref["status"]), f"{label}: {np.count_nonzero(st != ref['status'])} status mismatches"
    assert np.array_equal(iters, ref["iters"]), f"{label}: {np.count_nonzero(iters != ref['iters'])} iteration counts"
    h = st != 0
    err = np.abs(vals[h] - ref["out"][h]) / np.maximum(1.0, np.abs(ref["out"][h]))
    exact = np.all(vals[h] == ref["out"][h], axis=1)
    kinds = np.bincount(st, minlength=5)
    print(f"{label}: {h.sum()} hits (converged {kinds[2]}, penetrating {kinds[4]}), {exact.sum()} bit-identical, "
          f"max rel err {err.max(initial=0):.3e}", flush=True)
    assert err.max(initial=0) <= TOL, f"{label}: {(err > TOL).any(axis=1).sum()} hits beyond {TOL}"
    return h


def _opts_kw(max_toi=None, **kw):
    out = dict(kw)
    if max_toi is not None:
        out["max_time_of_impact"] = max_toi
    return out


def test_mixed_casts_match_oracle(orc):
    descs, verts, poses, vels, pairs = scenes.approach_pairs(100000, 7)
    shapes = ShapeSet(descs, verts)
    cases = [
        ("defaults", {}),
        ("target 0.1", dict(target_distance=0.1)),
        ("max_toi 1.5", dict(max_toi=1.5)),
        ("no stop at penetration", dict(stop_at_penetration=False)),
        ("no impact geometry", dict(compute_impact_geometry_on_penetration=False)),
    ]
    for label, kw in cases:
        print(f"case: casts, 100000 mixed pairs seed 7, {label}", flush=True)
        ref = orc.cast_shapes(descs, verts, poses, vels, pairs, **kw)
        h = _compare(label, _run(shapes, poses, vels, pairs, **_opts_kw(**kw)), ref)
        assert h.sum() > 10000


def test_penetrating_pairs_match_oracle(orc):
    # shapes that start overlapping: toi 0 and the impact geometry from GJK/EPA
    descs, verts, poses, vels, pairs = scenes.approach_pairs(60000, 9, spread=(0.0, 1.2))
    shapes = ShapeSet(descs, verts)
    for label, kw in (("defaults", {}), ("no stop at penetration", dict(stop_at_penetration=False)),
                      ("target 0.2, no stop", dict(target_distance=0.2, stop_at_penetration=False))):
        print(f"case: casts, 60000 overlapping pairs seed 9, {label}", flush=True)
        ref = orc.cast_shapes(descs, verts, poses, vels, pairs, **kw)
        got = _run(shapes, poses, vels, pairs, **kw)
        _compare(label, got, ref)
        assert (got[1] == 4).sum() > 10000


def test_ball_ball_and_reference_cases(orc):
    import torch

    n = 20000
    rng = np.random.default_rng(4)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["params"][:, 0] = rng.uniform(0.1, 1.0, 2 * n)
    poses = np.concatenate([scenes.random_rotations(2 * n, rng), rng.uniform(-3, 3, (2 * n, 3))], 1).astype(np.float32)
    vels = rng.uniform(-2, 2, (2 * n, 3)).astype(np.float32)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    shapes = ShapeSet(d, np.zeros((0, 3), np.float32))
    for label, kw in (("defaults", {}), ("target 0.3, no stop", dict(target_distance=0.3, stop_at_penetration=False))):
        print(f"case: ball-ball casts, {n} pairs, {label}", flush=True)
        _compare(f"ball-ball {label}", _run(shapes, poses, vels, pairs, **kw),
                 orc.cast_shapes(d, np.zeros((0, 3), np.float32), poses, vels, pairs, **kw))
    # the reference's tests (time_of_impact3.rs, ball_ball_toi.rs, ball_triangle_toi.rs, still_objects_toi.rs)
    print("case: reference shape-cast tests", flush=True)
    idq = [0, 0, 0, 1]
    tri = np.array([[0.5, -0.5, 0.0], [-0.5, -0.5, 0.0], [-0.5, 0.5, 0.0]], np.float32)
    kd = np.zeros(6, SHAPE_DESC_DTYPE)
    kd["type"] = [0, 1, 0, 0, 5, 0]
    kd["params"][0, 0] = 1.0
    kd["params"][1, :3] = 1.0
    kd["params"][2, 0] = 0.5
    kd["params"][3, 0] = 0.375
    kd["num_vertices"][4] = 3
    kd["params"][5, 0] = 0.5
    ks = ShapeSet(kd, tri)
    cases = [  # (pose1, vel1, shape1, pose2, vel2, shape2, expected toi or None)
        (idq + [1, 1, 1], [2, 2, 2], 0, idq + [0, 0, 0], [-1, 1, 1], 1, 0.0),
        (idq + [2, 2, 2], [-0.5, -0.5, -0.5], 0, idq + [0, 0, 0], [1, 1, 1], 1,
         (np.sqrt(np.float32(3.0)) - np.float32(1.0)) / np.linalg.norm(np.array([-1.5, -1.5, -1.5], np.float32))),
        (idq + [3, 3, 3], [2, 2, 2], 0, idq + [0, 0, 0], [-1, 1, 1], 1, None),
        (idq + [0, 0, 0], [0, 10, 0], 2, idq + [0, 10, 0], [0, 0, 0], 5, np.float32(0.9)),
        (idq + [0, 0, 0], [0, 6.925e-42, 0], 3, idq + [11.5, 5.5, 0], [0, 0, 0], 4, None),
    ]
    for c in cases:
        # one shape set; the case's two shapes at their poses and velocities
        p = np.zeros((6, 7), np.float32)
        p[:, 3] = 1.0
        v = np.zeros((6, 3), np.float32)
        i, j = c[2], c[5]
        p[i], v[i], p[j], v[j] = c[0], c[1], c[3], c[4]
        vals, st, _ = _run(ks, p, v, np.array([[i, j]], np.uint32))
        if c[6] is None:
            assert st[0] == 0, c
        else:
            assert st[0] != 0, c
            assert abs(vals[0, 0] - c[6]) <= 1e-6 * max(1.0, abs(c[6])), (c, vals[0, 0])
    # the stacked cuboids of still_objects_toi.rs
    sd = np.zeros(2, SHAPE_DESC_DTYPE)
    sd["type"] = 1
    sd["params"][:, :3] = 0.5
    ss = ShapeSet(sd, np.zeros((0, 3), np.float32))
    for vy, expect in ((0.0, None), (1.0, None), (-1.0, 0.1)):
        p = np.array([idq + [0, 1.1, 0], idq + [0, 0, 0]], np.float32)
        v = np.array([[0, vy, 0], [0, 0, 0]], np.float32)
        vals, st, _ = _run(ss, p, v, np.array([[0, 1]], np.uint32))
        assert (st[0] != 0) == (expect is not None)
        if expect is not None:
            assert abs(vals[0, 0] - expect) < 1e-5
    torch.cuda.synchronize()
