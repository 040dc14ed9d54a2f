"""Pins the oracle's broadphase restatement.

  * traverse_bvtt_single_tree (bvh_traverse_bvtt.rs:19-204) reports exactly the
    brute-force set of closed-overlapping leaf pairs, each once, on random,
    clustered, touching and identical boxes;
  * refit (bvh_refit.rs) after moving every leaf gives every internal box the
    exact min/max of its descendant leaves and keeps the depth-first layout;
  * the quaternion restatement of nalgebra rotates like the rotation matrix of
    the same unit quaternion (f64 closed form), inverts exactly on composition,
    and the ConvexPolyhedron AABB equals the min/max of the rotated points.
"""
import numpy as np
import pytest

from parry_amd import scenes


def random_boxes(n, seed, spread=20.0, size=2.0):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(0, spread, (n, 3))
    return np.concatenate([lo, lo + rng.uniform(0.1, size, (n, 3))], 1).astype(np.float32)


@pytest.mark.parametrize("case", ["random", "clustered", "touching", "identical", "tiny"])
def test_bvtt_equals_brute_force(orc, case):
    if case == "random":
        a = random_boxes(1500, 1)
    elif case == "clustered":
        a = np.concatenate([random_boxes(500, 2, 3.0), random_boxes(500, 3, 50.0)])
    elif case == "touching":  # faces exactly touching: closed intervals overlap
        x = np.arange(40, dtype=np.float32)
        g = np.stack(np.meshgrid(x, x, indexing="ij"), -1).reshape(-1, 2)
        a = np.concatenate([g, np.zeros((g.shape[0], 1)), g + 1, np.ones((g.shape[0], 1))], 1).astype(np.float32)
    elif case == "identical":
        a = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (60, 1))
    else:
        a = random_boxes(3, 4)
    bvh = orc.OracleBvh(a)
    pairs, touched = bvh.self_pairs()
    brute = orc.brute_pairs(a)
    assert len(orc.pair_set(pairs)) == pairs.shape[0]  # each pair once
    assert orc.pair_set(pairs) == orc.pair_set(brute)
    assert (pairs[:, 0] != pairs[:, 1]).all()
    if case == "touching":
        # king-move neighbours on a 40 x 40 grid: 2 * 39 * 40 straight + 2 * 39 * 39 diagonal
        assert brute.shape[0] == 2 * 39 * 40 + 2 * 39 * 39
    if case == "identical":
        assert pairs.shape[0] == 60 * 59 // 2


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_bvtt_tiny_trees(orc, n):
    a = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (n, 1))
    if n == 0:
        a = np.zeros((0, 6), np.float32)
    pairs, _ = orc.OracleBvh(a).self_pairs()
    assert pairs.shape[0] == n * (n - 1) // 2


def _check_boxes(nodes):
    lc = lambda d: int(d) & 0x3FFFFFFF
    for w in nodes:
        for side in ("left", "right"):
            if lc(w[side + "_data"]) > 1:
                ch = nodes[int(w[side + "_children"])]
                assert np.array_equal(w[side + "_mins"], np.minimum(ch["left_mins"], ch["right_mins"]))
                assert np.array_equal(w[side + "_maxs"], np.maximum(ch["left_maxs"], ch["right_maxs"]))


def test_refit_after_moving_leaves(orc):
    a = random_boxes(3000, 5)
    bvh = orc.OracleBvh(a)
    before = bvh.nodes()
    rng = np.random.default_rng(6)
    moved = (a.reshape(-1, 2, 3) + rng.uniform(-3, 3, (a.shape[0], 1, 3))).reshape(-1, 6).astype(np.float32)
    bvh.update_and_refit(moved)
    after = bvh.nodes()
    # topology unchanged, boxes exact
    for side in ("left", "right"):
        assert np.array_equal(before[side + "_children"], after[side + "_children"])
        assert np.array_equal(before[side + "_data"] & 0x3FFFFFFF, after[side + "_data"] & 0x3FFFFFFF)
    _check_boxes(after)
    # leaf slots hold the new boxes
    leaf = {}
    for w in after:
        for side in ("left", "right"):
            if (int(w[side + "_data"]) & 0x3FFFFFFF) == 1:
                leaf[int(w[side + "_children"])] = np.concatenate([w[side + "_mins"], w[side + "_maxs"]])
    assert all(np.array_equal(leaf[i], moved[i]) for i in range(a.shape[0]))
    # first refit turns pending flags into CHANGED, the next clears them (bvh_tree.rs:192-198)
    assert ((before["left_data"] >> 30) == 1).all()
    assert ((after["left_data"] >> 30) == 0).all()
    pairs, _ = bvh.self_pairs()
    assert orc.pair_set(pairs) == orc.pair_set(orc.brute_pairs(moved))


def quat_matrix(q):
    x, y, z, w = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def test_polyhedron_aabb_rotation_closed_form(orc):
    verts, offs = scenes.convex_polyhedra(200, seed=7)
    rng = np.random.default_rng(8)
    q = scenes.random_rotations(200, rng)
    t = rng.uniform(-10, 10, (200, 3))
    poses = np.concatenate([q, t], 1)
    a64 = orc.polyhedra_aabbs(verts, offs, poses, f64=True)
    a32 = orc.polyhedra_aabbs(verts, offs, poses.astype(np.float32))
    for i in range(200):
        p = verts[offs[i]:offs[i + 1]].astype(np.float64) @ quat_matrix(q[i]).T + t[i]
        ref = np.concatenate([p.min(0), p.max(0)])
        assert np.allclose(a64[i], ref, atol=1e-12)
        assert np.allclose(a32[i], ref, atol=2e-5)
    # identity pose: exact min/max of the vertices
    ident = np.tile(np.array([0, 0, 0, 1, 0, 0, 0], np.float32), (200, 1))
    a = orc.polyhedra_aabbs(verts, offs, ident)
    for i in range(200):
        p = verts[offs[i]:offs[i + 1]]
        assert np.array_equal(a[i], np.concatenate([p.min(0), p.max(0)]))
