"""Parity at the full sizes of the two contact workloads the bench measures.

  * learned warm start: 1,048,576 overlapping-AABB pairs (the first 2^20 of the
    GPU pair list of mixed_scene(320000, seed=3), 5 shape types uniform),
    prediction 0.01: query::contact with the network's seeds, and with parry's
    default seed, against the oracle given the same seeds (has_contact and GJK
    iterations identical, values within 1e-5 relative);
  * dynamic scene: mixed_scene(262144, seed=4) in rigid motion for 100 frames
    (learned seeds plus the per-pair cache of last frame's normals), with the
    pair set and every contact checked against the oracle on frames 0, 1, 2, 49
    and 99 (the poses those frames query, the seeds the GPU used).

These run the paths that only appear at scale: the EPA continuation-record pool
(one record per 48 pairs) and the long-pair tiers.
"""
import numpy as np
import pytest

from parry_amd import Bvh, ShapeSet, contact_pairs, scenes

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _canonical(pairs):
    p = np.sort(np.asarray(pairs, np.uint32).reshape(-1, 2), axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def _check_contacts(label, got, iters, ref):
    has = got.view(np.uint32)[:, 13]
    bad = np.count_nonzero(has.astype(np.int64) != ref["status"].astype(np.int64))
    assert bad == 0, f"{label}: {bad} has_contact mismatches"
    c = ref["status"] == 1
    err = np.abs(got[c, :13] - ref["out"][c]) / np.maximum(1.0, np.abs(ref["out"][c]))
    exact = int(np.all(got[c, :13] == ref["out"][c], axis=1).sum())
    print(f"{label}: {got.shape[0]} pairs, {int(c.sum())} contacts, {exact} bit-identical, "
          f"max rel err {err.max(initial=0):.3e}", flush=True)
    assert err.max(initial=0) <= TOL, f"{label}: contact error {err.max()}"
    if iters is not None:
        assert np.array_equal(iters[:, 0], ref["iters"][:, 0].astype(np.int32)), f"{label}: GJK iterations differ"


def test_learned_warm_start_full_size(orc):
    import torch

    from parry_amd.learned import WarmStartModel

    print("case: learned warm start, 1048576 pairs of mixed_scene(320000, 3)", flush=True)
    sc = scenes.mixed_scene(320000, seed=3)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    pairs = Bvh(shapes.compute_aabbs(poses).cpu().numpy()).self_pairs()
    assert pairs.shape[0] >= 1 << 20
    pairs = pairs[: 1 << 20].contiguous()
    net = WarmStartModel.pretrained()
    seeds = net.seeds(poses, pairs, shapes.descriptors())
    pr = pairs.cpu().numpy().view(np.uint32)
    for label, s in (("learned seeds", seeds), ("default seed", None)):
        out, iters = contact_pairs(shapes, poses, pairs, 0.01, init_dirs=s, with_iters=True)
        torch.cuda.synchronize()
        sd = None if s is None else s.cpu().numpy()
        ref = orc.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pr, 0.01, init_dirs=sd)
        _check_contacts(label, out.cpu().numpy(), iters.cpu().numpy(), ref)
        epa = ref["iters"][:, 1] > 0
        print(f"{label}: {int(epa.sum())} pairs through EPA, longest {int(ref['iters'][:, 1].max()) - 1} iterations",
              flush=True)


def test_dynamic_scene_full_size(orc):
    import torch

    from parry_amd.dynamics import DynamicScene

    print("case: dynamic scene, mixed_scene(262144, 4), 100 frames, learned+cache seeds", flush=True)
    sc = scenes.mixed_scene(262144, seed=4)
    descs, verts = sc["descs"], sc["vertices"]
    ds = DynamicScene(descs, verts, sc["poses"], seed=104, seeding="learned+cache",
                      pair_capacity=4 * 262144 + (1 << 20))
    checked = (0, 1, 2, 49, 99)
    for f in range(100):
        poses = None
        if f in checked:
            torch.cuda.synchronize()  # the previous frame's pose update runs on the scene's stream
            poses = ds.poses.cpu().numpy().copy()  # the poses this frame queries
        ds.step(use_graph=False)
        if f not in checked:
            continue
        m = ds.num_pairs()
        torch.cuda.synchronize()
        pairs = ds.pairs[:m].cpu().numpy().view(np.uint32).copy()
        contacts = ds.contacts[:m].cpu().numpy().copy()
        seeds = ds.seeds[:m].cpu().numpy().copy()
        print(f"case: dynamic frame {f}, {m} pairs", flush=True)
        ref_pairs, _ = orc.OracleBvh(orc.shape_aabbs(descs, verts, poses)).self_pairs()
        assert np.array_equal(_canonical(pairs), _canonical(ref_pairs)), f"frame {f}: pair sets differ"
        ref = orc.contact_pairs(descs, verts, poses, pairs, 0.01, init_dirs=seeds)
        _check_contacts(f"frame {f}", contacts, None, ref)
