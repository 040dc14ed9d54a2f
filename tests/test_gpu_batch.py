"""GPU parity of the batched multi-mesh ray cast (one launch over many TriMeshes).

Bar: every ray's record (feature, time of impact, normal) is byte-identical to
the same ray cast alone on its own mesh by the single-mesh kernel, and to the
f32 oracle (parry3d's TriMesh::cast_local_ray_and_get_normal, per mesh), for
both lane mappings, after autotuning, and on every replay of a captured graph.
"""
import numpy as np
import pytest

from parry_amd import RAY_HIT_DTYPE, TriMesh, TriMeshBatch, scenes

pytestmark = pytest.mark.gpu

NO_HIT = 0xFFFFFFFF
MAX_TOI = 10.0


def _torch():
    import torch

    return torch


def _as_records(t):
    return t.cpu().numpy().view(RAY_HIT_DTYPE).reshape(-1)


def _build(spec):
    meshes, rays = [], []
    for level, n, mseed, rseed in spec:
        if level < 0:  # one-triangle mesh
            v = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], dtype=np.float32)
            f = np.array([[0, 1, 2]], dtype=np.uint32)
        else:
            v, f = scenes.noisy_icosphere(level, seed=mseed)
        meshes.append((v, f, TriMesh(v, f)))
        rays.append(scenes.rays_into_unit_ball(n, seed=rseed))
    return meshes, rays


SMALL = [  # (level, rays, mesh seed, ray seed); -1 = single triangle, 0 rays = empty range
    (2, 700, 11, 12), (0, 0, 13, 14), (4, 3000, 15, 16), (-1, 257, 0, 18), (1, 64, 19, 20),
    (3, 1, 21, 22), (0, 5000, 23, 24), (4, 0, 25, 26), (2, 129, 27, 28), (3, 4096, 29, 30),
]


# Kernel mappings of the batched launch (pq_ray_tuning): results must not depend on them.
BATCH_MAPPINGS = (
    [{"mapping": "static", "block_size": b, "stack_regs": r} for b, r in ((256, 0), (64, 1), (128, 2), (256, 4))]
    + [{"mapping": "persistent", "chunk": c} for c in (64, 128, 512)]
    + [{"mapping": "persistent", "chunk": 128, "block_size": 64, "stack_regs": 2},
       {"mapping": "wave", "chunk": 1}, {"mapping": "wave", "chunk": 4}]
    + [{"mapping": "binned", "block_size": 64}, {"mapping": "binned", "block_size": 256, "stack_regs": 2}]
)


def _check_batch(orc, meshes, rays, mappings=("static", "persistent", "wave", "binned")):
    torch = _torch()
    batch = TriMeshBatch([m for _, _, m in meshes], [r.shape[0] for r in rays])
    all_rays = np.concatenate(rays) if rays else np.zeros((0, 6), np.float32)
    d_rays = torch.from_numpy(all_rays).cuda()
    # per-mesh single-kernel results and the oracle
    single, ref = [], []
    for (v, f, mesh), r in zip(meshes, rays):
        mesh.set_tuning("static")
        single.append(mesh.cast_local_rays_and_get_normal(r, MAX_TOI, True) if r.shape[0] else
                      np.zeros(0, RAY_HIT_DTYPE))
        ref.append(orc.OracleTriMesh(v, f).cast_rays(r, MAX_TOI) if r.shape[0] else None)
    single = np.concatenate(single)
    for tu in [t for t in BATCH_MAPPINGS if t["mapping"] in mappings]:
        label = "/".join(f"{k}={v}" for k, v in tu.items())
        print(f"case: batch of {len(meshes)} meshes, {all_rays.shape[0]} rays, {label}", flush=True)
        batch.set_tuning(**tu)
        hits = _as_records(batch.cast_local_rays_and_get_normal(d_rays, MAX_TOI, True))
        assert hits.tobytes() == single.tobytes(), f"{label}: batch differs from per-mesh casts"
        for i, rr in enumerate(ref):
            if rr is None:
                continue
            lo, hi = int(batch.offsets[i]), int(batch.offsets[i + 1])
            h = hits[lo:hi]
            assert np.array_equal(h["feature"], rr["feature"]), f"mesh {i}: feature mismatch"
            hit = h["feature"] != NO_HIT
            assert np.array_equal(h["time_of_impact"][hit], rr["toi"][hit]), f"mesh {i}: toi mismatch"
            assert np.array_equal(h["normal"][hit], rr["normal"][hit]), f"mesh {i}: normal mismatch"
    return batch, d_rays, single


def test_small_uneven_batch(orc):
    meshes, rays = _build(SMALL)
    _check_batch(orc, meshes, rays)


def test_batch_with_axis_aligned_rays(orc):
    # Every 7th ray along a coordinate axis and a few with a zero direction,
    # scattered so that most waves mix them with ordinary rays (the kernels
    # take the degenerate-axis slab test for such waves).
    meshes, rays = _build(SMALL)
    for k, r in enumerate(rays):
        idx = np.arange(k % 7, r.shape[0], 7)
        r[idx, 3:] = 0.0
        r[idx, 3 + idx % 3] = np.where(idx % 2 == 0, 1.0, -1.0)
        r[idx, :3] *= 0.5
        r[idx[::50], 3:] = 0.0
    _check_batch(orc, meshes, rays)


def test_counted_batch_matches_single_mesh_counts():
    torch = _torch()
    meshes, rays = _build(SMALL)
    batch = TriMeshBatch([m for _, _, m in meshes], [r.shape[0] for r in rays])
    d_rays = torch.from_numpy(np.concatenate(rays)).cuda()
    for tu in BATCH_MAPPINGS:
        batch.set_tuning(**tu)
        mapping = "/".join(f"{k}={v}" for k, v in tu.items())
        hits, counts = batch.cast_local_rays_counted(d_rays, MAX_TOI)
        counts = counts.cpu().numpy()
        for i, ((_, _, mesh), r) in enumerate(zip(meshes, rays)):
            if r.shape[0] == 0:
                continue
            mesh.set_tuning("static")
            _, c1 = mesh.cast_local_rays_counted(torch.from_numpy(r).cuda(), MAX_TOI)
            lo, hi = int(batch.offsets[i]), int(batch.offsets[i + 1])
            assert np.array_equal(counts[lo:hi], c1.cpu().numpy()), f"mesh {i} ({mapping}): traversal counts differ"


def test_autotune_and_graph_replays(orc):
    torch = _torch()
    meshes, rays = _build(SMALL)
    batch, d_rays, single = _check_batch(orc, meshes, rays, mappings=("static",))
    out = torch.empty((batch.num_rays, 5), dtype=torch.float32, device="cuda")
    tuning, ms = batch.autotune(d_rays, MAX_TOI, out=out)
    print(f"autotune chose {tuning} at {ms:.4f} ms", flush=True)
    assert tuning["mapping"] in ("static", "persistent", "wave", "binned") and ms > 0
    assert _as_records(out).tobytes() == single.tobytes()
    out.zero_()
    graph = batch.capture(d_rays, MAX_TOI, True, out)
    for rep in range(4):
        out.fill_(float(rep + 1))  # eager work between replays
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert _as_records(out).tobytes() == single.tobytes(), f"graph replay {rep} differs"
    # a replay after the rays change in place sees the new rays
    perm = torch.randperm(d_rays.shape[0], device="cuda")
    eager = batch.cast_local_rays_and_get_normal(d_rays[perm].contiguous(), MAX_TOI, True)
    d_rays.copy_(d_rays[perm].contiguous())
    graph.launch()
    torch.cuda.synchronize()
    # compare bit patterns: a miss stores feature 0xFFFFFFFF, a NaN when read as f32
    assert torch.equal(out.view(torch.int32), eager.view(torch.int32))


def test_binned_batch_graph_replays(orc):
    """The binned mapping of a batch (a sort per mesh into the batch's own
    buffers, then the static kernel over the sorted order) under graph capture:
    every replay equals the per-mesh casts, and a replay after the rays change
    in place sees the new rays."""
    torch = _torch()
    meshes, rays = _build(SMALL)
    batch, d_rays, single = _check_batch(orc, meshes, rays, mappings=("binned",))
    out = torch.empty((batch.num_rays, 5), dtype=torch.float32, device="cuda")
    batch.set_tuning("binned", block_size=128)
    graph = batch.capture(d_rays, MAX_TOI, True, out)
    for rep in range(4):
        out.fill_(float(rep + 1))
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert _as_records(out).tobytes() == single.tobytes(), f"binned graph replay {rep} differs"
    perm = torch.randperm(d_rays.shape[0], device="cuda")
    d_rays.copy_(d_rays[perm].contiguous())
    batch.set_tuning("static")
    eager = batch.cast_local_rays_and_get_normal(d_rays, MAX_TOI, True)
    graph.launch()
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), eager.view(torch.int32))


def test_invalid_batches_are_rejected():
    from parry_amd import _native

    v, f = scenes.noisy_icosphere(1, seed=1)
    m = TriMesh(v, f)
    with pytest.raises(_native.NativeError):
        TriMeshBatch([], [])
    with pytest.raises(ValueError):
        TriMeshBatch([m], [1, 2])


def test_uneven_64_mesh_workload(orc):
    """The benchmark workload: 64 meshes at levels 3-7 with 2^10-2^20 rays each."""
    spec = scenes.uneven_mesh_batch(64, seed=2)
    meshes, rays = _build(spec)
    _check_batch(orc, meshes, rays)
