"""Pins how the oracle's EPA (epa3.rs:428-651) ends, and what each ending returns.

EPA::closest_points leaves its loop in one of these ways (oracle.EPA_EXIT_NAMES):

  converged   max_dist - curr_dist < 100 EPS (epa3.rs:568): returns the *best*
              face, the one whose support point gave the smallest upper bound;
  stuck       |curr_dist - old_dist| < EPS && candidate_max_dist < max_dist
              (epa3.rs:569-572).  It cannot hold: max_dist was just lowered to
              at most candidate_max_dist (epa3.rs:560-563);
  numerical   a new face closer to the origin than the popped one
              (epa3.rs:614-621): returns the popped face;
  iter_cap    more than 100 iterations (epa3.rs:640-645): returns the best face;
  heap_empty  every face popped: returns the best face.

For polytopes (Cuboid, ConvexPolyhedron) the exact penetration depth is known
independently: the minimum over the Minkowski difference's face normals (faces of
A, negated faces of B, edge-edge cross products) of the support of A - B
(separating-axis theorem, float64).  The checks:

  * no pair ever leaves through the stuck clause;
  * the normal of a converged exit is the minimum-translation axis: the support
    of A - B along it equals the SAT depth;
  * no exit reports a depth beyond the SAT depth;
  * the reported depth is the *plane distance of the best face*, which can be
    shallower than the SAT depth when that face was popped before the polytope
    reached the boundary along its normal.  The path to the best face differs
    between f32 and f64, so the two precisions can report different depths with
    the same normal: this is the source of the large f32-vs-f64 differences in
    the dynamic-scene error report, not rounding in either one.
"""
import numpy as np
from scipy.spatial import ConvexHull
from scipy.spatial.transform import Rotation

from parry_amd import scenes

SH_CUBOID, SH_POLY = 1, 4


def _world_points(descs, verts, poses, i):
    d = descs[i]
    if d["type"] == SH_POLY:
        p = verts[d["first_vertex"]:d["first_vertex"] + d["num_vertices"]].astype(np.float64)
    else:
        he = d["params"][:3].astype(np.float64)
        p = np.array([[sx * he[0], sy * he[1], sz * he[2]] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    R = Rotation.from_quat(poses[i, :4].astype(np.float64)).as_matrix()
    return p @ R.T + poses[i, 4:].astype(np.float64)


def _faces_and_edges(P):
    h = ConvexHull(P)
    e = set()
    for s in h.simplices:
        for a, b in ((s[0], s[1]), (s[1], s[2]), (s[2], s[0])):
            e.add((min(a, b), max(a, b)))
    e = np.array(sorted(e))
    return h.equations[:, :3], P[e[:, 1]] - P[e[:, 0]]


def sat_depth(A, B):
    """Minimum translation distance of overlapping polytopes A, B (float64)."""
    na, ea = _faces_and_edges(A)
    nb, eb = _faces_and_edges(B)
    cr = np.cross(ea[:, None, :], eb[None, :, :]).reshape(-1, 3)
    ln = np.linalg.norm(cr, axis=1)
    cr = cr[ln > 1e-9] / ln[ln > 1e-9, None]
    D = np.concatenate([na, -nb, cr, -cr])
    return float(((A @ D.T).max(0) - (B @ D.T).min(0)).min())


def support_gap(A, B, n):
    """Support of A - B along the unit normal n: how far B must move along n."""
    return float((A @ n).max() - (B @ n).min())


def _scene(orc):
    sc = scenes.mixed_scene(16384, seed=4)
    aabbs = orc.shape_aabbs(sc["descs"], sc["vertices"], sc["poses"])
    pairs, _ = orc.OracleBvh(aabbs).self_pairs()
    r32 = orc.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pairs, 0.01)
    r64 = orc.contact_pairs(sc["descs"], sc["vertices"].astype(np.float64), sc["poses"].astype(np.float64), pairs,
                            0.01, f64=True)
    return sc, pairs, r32, r64


def test_epa_exits_against_sat(orc):
    names = orc.EPA_EXIT_NAMES
    sc, pairs, r32, r64 = _scene(orc)
    stuck = names.index("stuck")
    # every shape type, both precisions: the stuck clause never fires
    assert (r32["epa_exit"] != stuck).all() and (r64["epa_exit"] != stuck).all()
    assert ((r32["epa_exit"] > 0) == (r32["iters"][:, 1] > 0)).all()

    types = sc["descs"]["type"]
    poly = np.isin(types[pairs[:, 0]], [SH_CUBOID, SH_POLY]) & np.isin(types[pairs[:, 1]], [SH_CUBOID, SH_POLY])
    epa = poly & (r32["status"] == 1) & (r64["status"] == 1) & (r32["epa_exit"] > 0) & (r64["epa_exit"] > 0)
    idx = np.flatnonzero(epa)
    assert idx.size > 1000
    # every pair whose two precisions disagree, plus a seeded sample of the rest
    differ = idx[np.abs(r32["out"][idx, 12] - r64["out"][idx, 12]) > 1e-4]
    rest = np.setdiff1d(idx, differ)
    sample = np.concatenate([differ, np.random.default_rng(0).choice(rest, 300, replace=False)])
    poses64 = sc["poses"].astype(np.float64)
    conv = names.index("converged")
    n_conv = n_exact = 0
    for k in sample:
        A = _world_points(sc["descs"], sc["vertices"], poses64, pairs[k, 0])
        B = _world_points(sc["descs"], sc["vertices"], poses64, pairs[k, 1])
        sat = sat_depth(A, B)
        for r, tol in ((r64, 1e-6), (r32, 2e-4)):
            depth = -float(r["out"][k, 12])
            n1 = r["out"][k, 6:9].astype(np.float64)
            gap = support_gap(A, B, n1 / np.linalg.norm(n1))
            assert depth <= sat + tol * max(1.0, sat), (k, depth, sat)  # never deeper than the true depth
            assert gap >= sat - tol * max(1.0, sat)  # SAT is the minimum over all axes
            if r["epa_exit"][k] == conv:
                assert abs(gap - sat) <= tol * max(1.0, sat), (k, gap, sat)  # the optimal axis
                if r is r64:
                    n_conv += 1
                    if abs(depth - sat) <= 1e-6 * max(1.0, sat):
                        n_exact += 1
    # most converged pairs report the exact depth
    assert n_exact >= 0.9 * n_conv, (n_exact, n_conv)


def test_converged_exit_reports_best_face_plane(orc):
    """Four pairs of the dynamic scene's first frame (tests/golden/
    make_epa_shallow_pairs.py) on which one precision's converged exit reports
    the best face's plane distance, shallower than the exact depth, while the
    other reports the exact depth; both normals are the optimal axis."""
    import os

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "epa_shallow_pairs.npz"))
    descs, verts, poses = z["descs"], z["vertices"], z["poses"]
    pairs = np.arange(descs.shape[0], dtype=np.uint32).reshape(-1, 2)
    r32 = orc.contact_pairs(descs, verts, poses, pairs, 0.01)
    r64 = orc.contact_pairs(descs, verts.astype(np.float64), poses.astype(np.float64), pairs, 0.01, f64=True)
    shallow_in = ("f32", "f32", "f64", "f64")
    conv = orc.EPA_EXIT_NAMES.index("converged")
    poses64 = poses.astype(np.float64)
    for k in range(pairs.shape[0]):
        A = _world_points(descs, verts, poses64, pairs[k, 0])
        B = _world_points(descs, verts, poses64, pairs[k, 1])
        sat = sat_depth(A, B)
        # (the stored f32 quaternions are unit only to ~1e-7, and the shapes sit ~75 from the origin)
        for name, r, tol in (("f32", r32, 2e-5), ("f64", r64, 1e-6)):
            assert r["status"][k] == 1 and r["epa_exit"][k] == conv
            n1 = r["out"][k, 6:9].astype(np.float64)
            assert abs(support_gap(A, B, n1 / np.linalg.norm(n1)) - sat) <= tol
            depth = -float(r["out"][k, 12])
            if name == shallow_in[k]:
                assert depth < sat - 1e-2, (k, name, depth, sat)
            else:
                assert abs(depth - sat) <= tol, (k, name, depth, sat)


def test_epa_best_face_plane_known_answer(orc):
    """Cuboid(2, 1, 1) against itself (the reference's KAT, epa3.rs:6-23 of the
    tests): EPA converges to the exact depth with the face normal, so converged
    exits are not shallow by construction."""
    from parry_amd import SHAPE_DESC_DTYPE

    descs = np.zeros(2, SHAPE_DESC_DTYPE)
    descs["type"] = SH_CUBOID
    descs["params"][:, :3] = (2, 1, 1)
    poses = np.array([[0, 0, 0, 1, 3.5, 0, 0], [0, 0, 0, 1, 0, 0, 0]], np.float32)
    r = orc.contact_pairs(descs, np.zeros((0, 3), np.float32), poses, [[0, 1]], 10.0)
    assert r["status"][0] == 1 and r["out"][0, 12] == np.float32(-0.5)
    assert orc.EPA_EXIT_NAMES[r["epa_exit"][0]] == "converged"
