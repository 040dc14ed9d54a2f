"""GPU parity of the broadphase: shape AABBs, Bvh build/refit and self pairs.

Bar (integer/bit work): leaf AABBs bit-exact with the f32 oracle, the device
tree after every refit byte-identical to the oracle's (boxes, leaf counts,
change flags), and the pair set equal to parry's traverse_bvtt_single_tree set
(the oracle's), which tests/test_oracle_broadphase.py pins to brute force.
"""
import numpy as np
import pytest

from parry_amd import Bvh, ShapeSet, scenes

pytestmark = pytest.mark.gpu


def _move(poses, seed, step=0.5, angle=0.3):
    """Small random rigid motion of every pose (f64 math, rounded to f32)."""
    rng = np.random.default_rng(seed)
    n = poses.shape[0]
    axis = rng.standard_normal((n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    th = rng.uniform(-angle, angle, n)
    dq = np.concatenate([axis * np.sin(th / 2)[:, None], np.cos(th / 2)[:, None]], 1)
    q = poses[:, :4].astype(np.float64)
    x1, y1, z1, w1 = dq.T
    x2, y2, z2, w2 = q.T
    nq = np.stack(
        [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
         w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
    nq /= np.linalg.norm(nq, axis=1, keepdims=True)
    t = poses[:, 4:].astype(np.float64) + rng.uniform(-step, step, (n, 3))
    return np.ascontiguousarray(np.concatenate([nq, t], 1), dtype=np.float32)


def test_c1_full_size(orc):
    """65,536 ConvexPolyhedra (the broadphase + contact workload), then one moved frame."""
    import torch

    print("case: C1 polyhedra scene n=65536 seed=3", flush=True)
    sc = scenes.polyhedra_scene(65536, seed=3)
    shapes = ShapeSet.convex_polyhedra(sc["vertices"], sc["offsets"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    aabbs = shapes.compute_aabbs(poses)
    torch.cuda.synchronize()
    ref_aabbs = orc.polyhedra_aabbs(sc["vertices"], sc["offsets"], sc["poses"])
    assert np.array_equal(aabbs.cpu().numpy(), ref_aabbs)

    bvh = Bvh(aabbs.cpu().numpy())
    obvh = orc.OracleBvh(ref_aabbs)
    assert bvh.nodes().tobytes() == obvh.nodes().tobytes()
    pairs = bvh.self_pairs()
    ref_pairs, touched = obvh.self_pairs()
    got = orc.pair_set(pairs.cpu().numpy())
    assert got == orc.pair_set(ref_pairs)
    assert len(got) == pairs.shape[0]
    mean_overlaps = 2 * len(got) / 65536
    print(f"C1 frame 0: {len(got)} pairs, mean AABB overlaps per shape {mean_overlaps:.3f}, "
          f"oracle BVTT wide records touched {touched}")
    assert 6.0 < mean_overlaps < 10.0

    # A moved frame: leaf update + refit, then pairs again.
    poses2 = _move(sc["poses"], seed=4)
    p2 = torch.from_numpy(poses2).cuda()
    aabbs2 = shapes.compute_aabbs(p2)
    bvh.update_and_refit(aabbs2)
    torch.cuda.synchronize()
    ref2 = orc.polyhedra_aabbs(sc["vertices"], sc["offsets"], poses2)
    assert np.array_equal(aabbs2.cpu().numpy(), ref2)
    obvh.update_and_refit(ref2)
    assert bvh.nodes().tobytes() == obvh.nodes().tobytes()
    pairs2 = bvh.self_pairs()
    assert orc.pair_set(pairs2.cpu().numpy()) == orc.pair_set(obvh.self_pairs()[0])
    # and a second refit (change flags clear on the second pass)
    poses3 = _move(poses2, seed=5)
    aabbs3 = shapes.compute_aabbs(torch.from_numpy(poses3).cuda())
    bvh.update_and_refit(aabbs3)
    obvh.update_and_refit(orc.polyhedra_aabbs(sc["vertices"], sc["offsets"], poses3))
    assert bvh.nodes().tobytes() == obvh.nodes().tobytes()


@pytest.mark.parametrize("case", ["empty", "one", "two", "three", "identical", "touching", "random"])
def test_small_and_degenerate(orc, case):
    import torch

    rng = np.random.default_rng(9)
    if case == "empty":
        a = np.zeros((0, 6), np.float32)
    elif case in ("one", "two", "three"):
        n = {"one": 1, "two": 2, "three": 3}[case]
        a = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (n, 1))
    elif case == "identical":
        a = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (70, 1))
    elif case == "touching":
        x = np.arange(30, dtype=np.float32)
        g = np.stack(np.meshgrid(x, x, indexing="ij"), -1).reshape(-1, 2)
        a = np.concatenate([g, np.zeros((g.shape[0], 1)), g + 1, np.ones((g.shape[0], 1))], 1).astype(np.float32)
    else:
        lo = rng.uniform(0, 30, (4000, 3))
        a = np.concatenate([lo, lo + rng.uniform(0.1, 2, (4000, 3))], 1).astype(np.float32)
    print(f"case: broadphase {case} n={a.shape[0]}", flush=True)
    bvh = Bvh(a)
    ob = orc.OracleBvh(a)
    assert bvh.nodes().tobytes() == ob.nodes().tobytes()
    pairs = bvh.self_pairs(capacity=3)  # forces the overflow + retry path when there are more pairs
    assert orc.pair_set(pairs.cpu().numpy()) == orc.pair_set(orc.brute_pairs(a))
    if a.shape[0]:
        moved = (a.reshape(-1, 2, 3) + rng.uniform(-1, 1, (a.shape[0], 1, 3))).reshape(-1, 6).astype(np.float32)
        bvh.update_and_refit(torch.from_numpy(moved).cuda())
        ob.update_and_refit(moved)
        assert bvh.nodes().tobytes() == ob.nodes().tobytes()
        assert orc.pair_set(bvh.self_pairs().cpu().numpy()) == orc.pair_set(orc.brute_pairs(moved))
