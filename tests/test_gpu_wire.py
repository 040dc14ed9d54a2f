"""Meshes and broadphase trees rebuilt from parry's serialized form run exactly as
the originals: the same ray hits (C ABI, bit for bit) and the same pair sets."""
import numpy as np
import pytest

from parry_amd import Bvh, TriMesh, scenes

pytestmark = pytest.mark.gpu


def test_trimesh_from_bytes_casts_identically():
    import torch

    v, f = scenes.noisy_icosphere(5, seed=3)
    mesh = TriMesh(v, f)
    data = mesh.to_bytes()
    copy = TriMesh.from_bytes(data)
    assert copy.bvh_nodes().tobytes() == mesh.bvh_nodes().tobytes()
    assert copy.to_bytes() == data
    rays = torch.from_numpy(scenes.rays_into_unit_ball(100000, seed=4)).cuda()
    print(f"case: TriMesh s=5 from {len(data)} serialized bytes, 100000 rays", flush=True)
    a = mesh.cast_local_rays_and_get_normal(rays, 10.0, True)
    b = copy.cast_local_rays_and_get_normal(rays, 10.0, True)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_bvh_from_bytes_pairs_identically():
    import torch

    from parry_amd import ShapeSet

    sc = scenes.mixed_scene(5000, seed=8)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    aabbs = shapes.compute_aabbs(poses)
    bvh = Bvh(aabbs.cpu().numpy())
    data = bvh.to_bytes()
    copy = Bvh.from_bytes(data)
    print(f"case: broadphase Bvh of 5000 shapes from {len(data)} serialized bytes", flush=True)
    assert copy.num_leaves == bvh.num_leaves and copy.nodes().tobytes() == bvh.nodes().tobytes()
    p0 = bvh.self_pairs().cpu().numpy()
    p1 = copy.self_pairs().cpu().numpy()
    assert np.array_equal(p0, p1)
    # and after a refit of both from moved shapes
    moved = aabbs + 0.01
    bvh.update_and_refit(moved)
    copy.update_and_refit(moved)
    assert np.array_equal(bvh.self_pairs().cpu().numpy(), copy.self_pairs().cpu().numpy())
