"""The dynamic scene (rigid motion, one collision frame per step) on the GPU.

  * every frame's pair set equals the oracle's (closed-interval AABB overlaps of
    the frame's poses; the GPU refits the frame-0 tree, the oracle builds anew);
  * every frame's contacts equal the oracle's query::contact given the same
    seeds (has_contact identical, values within 1e-5);
  * frames replayed from the captured hipGraphs are byte-identical to frames
    issued eagerly;
  * the warm-start cache carries normals of persisting pairs to the next frame.
"""
import numpy as np
import pytest

from parry_amd import scenes
from parry_amd.dynamics import DynamicScene

pytestmark = pytest.mark.gpu


def _scene(n, seed, seeding="learned+cache"):
    sc = scenes.mixed_scene(n, seed)
    return sc, DynamicScene(sc["descs"], sc["vertices"], sc["poses"], seed=seed + 100, seeding=seeding)


def _snapshot(ds):
    import torch

    torch.cuda.synchronize()
    m = ds.num_pairs()
    return {
        "pairs": ds.pairs[:m].cpu().numpy().view(np.uint32).copy(),
        "contacts": ds.contacts[:m].cpu().numpy().copy(),
        "seeds": ds.seeds[:m].cpu().numpy().copy(),
        "poses": ds.poses.cpu().numpy().copy(),
    }


def test_frames_match_oracle(orc):
    import torch

    sc, ds = _scene(8000, 31)
    descs, verts = sc["descs"], sc["vertices"]
    for f in range(5):
        torch.cuda.synchronize()
        poses = ds.poses.cpu().numpy().copy()  # the poses this frame queries
        ds.step(use_graph=False)
        snap = _snapshot(ds)
        print(f"case: dynamic frame {f}, {snap['pairs'].shape[0]} pairs", flush=True)
        aabbs = orc.shape_aabbs(descs, verts, poses)
        ref_pairs, _ = orc.OracleBvh(aabbs).self_pairs()
        assert orc.pair_set(snap["pairs"]) == orc.pair_set(ref_pairs), f"frame {f}: pair sets differ"
        ref = orc.contact_pairs(descs, verts, poses, snap["pairs"], 0.01, init_dirs=snap["seeds"])
        has = snap["contacts"].view(np.uint32)[:, 13]
        assert np.array_equal(has.astype(np.int64), ref["status"].astype(np.int64)), f"frame {f}: has_contact"
        c = ref["status"] == 1
        err = np.abs(snap["contacts"][c, :13] - ref["out"][c]) / np.maximum(1.0, np.abs(ref["out"][c]))
        assert err.max(initial=0) <= 1e-5, f"frame {f}: contact error {err.max()}"


def test_graph_replay_matches_eager():
    _, eager = _scene(6000, 41)
    _, graph = _scene(6000, 41)
    graph.capture()
    for f in range(6):
        eager.step(use_graph=False)
        graph.step(use_graph=True)
        a, b = _snapshot(eager), _snapshot(graph)
        print(f"case: graph replay frame {f}, {a['pairs'].shape[0]} pairs", flush=True)
        for k in ("pairs", "contacts", "seeds", "poses"):
            assert a[k].tobytes() == b[k].tobytes(), f"frame {f}: {k} differ between graph and eager"


def test_warm_cache_carries_normals():
    import ctypes

    import torch

    from parry_amd import _native

    sc, ds = _scene(6000, 51, seeding="cache")
    ds.step(use_graph=False)
    first = _snapshot(ds)
    # frame 1 reads the cache frame 0 wrote (cache[1])
    hits = torch.zeros(1, dtype=torch.int32, device="cuda")
    seeds = torch.zeros_like(ds.seeds)
    torch.cuda.synchronize()
    _native.check(_native.lib().pq_warm_cache_apply(ds.caches[1].data_ptr(), ds.cache_cap, ds.pairs.data_ptr(),
                                                    ds.count.data_ptr(), ds.capacity, seeds.data_ptr(),
                                                    hits.data_ptr(), None))
    torch.cuda.synchronize()
    n_contacts = int((first["contacts"].view(np.uint32)[:, 13] == 1).sum())
    print(f"cache: {n_contacts} contacts stored, {int(hits.item())} pairs found")
    assert int(hits.item()) == n_contacts  # same pair list: every stored pair is found
    m = first["pairs"].shape[0]
    got = seeds[:m].cpu().numpy()
    has = first["contacts"].view(np.uint32)[:, 13] == 1
    nrm = np.linalg.norm(got[has], axis=1)
    assert np.allclose(nrm, 1.0, atol=1e-5) and (np.abs(got[~has]) == 0).all()
