"""The flop-counting instantiation of the oracle's query::contact (oracle_flops.cpp),
which the bench divides by the GPU's contact time for the FP32 fraction.

  * it computes exactly what the f32 restatement computes (same statuses, the
    same dist bit for bit), so it counts the operations of that computation;
  * the per-pair counts add up to the totals;
  * the closed-form Ball-Ball branch costs the same on every pair, and a
    polyhedron's support scan (3 mul + 2 add per vertex per support call) makes
    the count grow with the vertex count.
"""
import numpy as np

from parry_amd import SHAPE_DESC_DTYPE, scenes


def test_counts_follow_the_f32_restatement(orc):
    sc = scenes.mixed_scene(4096, seed=7)
    pairs, _ = orc.OracleBvh(orc.shape_aabbs(sc["descs"], sc["vertices"], sc["poses"])).self_pairs()
    f = orc.contact_flops(sc["descs"], sc["vertices"], sc["poses"], pairs, 0.01)
    r = orc.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pairs, 0.01)
    assert np.array_equal(f["status"], r["status"])
    c = r["status"] == 1
    assert np.array_equal(f["dist"][c], r["out"][c, 12])
    assert int(f["per_pair"].astype(np.uint64).sum()) == f["flops"]
    assert f["flops"] == sum(f["counts"][k] for k in ("add", "mul", "div", "sqrt"))


def test_ball_ball_constant_and_polyhedron_scaling(orc):
    rng = np.random.default_rng(3)
    n = 64
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["params"][:, 0] = rng.uniform(0.5, 1.0, 2 * n)
    poses = np.zeros((2 * n, 7), np.float32)
    poses[:, 3] = 1
    poses[1::2, 4:] = rng.normal(0, 0.5, (n, 3))
    pairs = np.arange(2 * n, dtype=np.uint32).reshape(-1, 2)
    f = orc.contact_flops(d, np.zeros((0, 3), np.float32), poses, pairs, 10.0)
    assert (f["status"] == 1).all() and len(set(f["per_pair"].tolist())) == 1
    # a separated pair of polyhedra, then the same pair with shape 1's vertices
    # listed twice: no support point changes, and each support call on shape 1
    # scans k more vertices at one dot product (3 mul + 2 add) each
    for k in (8, 16, 32, 64):
        pts = scenes.convex_hull_vertices(scenes.convex_point_cloud(np.random.default_rng(k), k)).astype(np.float32)
        nv = pts.shape[0]
        verts = np.concatenate([pts, pts, pts])  # shape 1: [0, nv) or [0, 2 nv); shape 2: [2 nv, 3 nv)
        dd = np.zeros(2, SHAPE_DESC_DTYPE)
        dd["type"] = 4
        dd["first_vertex"] = (0, 2 * nv)
        dd["num_vertices"] = (nv, nv)
        p2 = np.array([[0, 0, 0, 1, 0, 0, 0], [0.2, 0.1, 0, 0.97467943, 5, 0.5, 0]], np.float32)
        base = orc.contact_flops(dd, verts, p2, [[0, 1]], 10.0)
        ref = orc.contact_pairs(dd, verts, p2, [[0, 1]], 10.0)
        dd["num_vertices"] = (2 * nv, nv)
        dup = orc.contact_flops(dd, verts, p2, [[0, 1]], 10.0)
        assert dup["status"][0] == base["status"][0] == 1 and dup["dist"][0] == base["dist"][0]
        extra = int(dup["per_pair"][0]) - int(base["per_pair"][0])
        assert extra > 0 and extra % (5 * nv) == 0
        # simplex.reset's support call, one per completed GJK iteration, and the last
        # (exiting) iteration's
        assert extra // (5 * nv) == int(ref["iters"][0, 0]) + 2
