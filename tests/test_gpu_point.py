"""GPU parity of the TriMesh point projection and of the culling ray cast (C ABI).

Bar: bit-identical to the f32 oracle -- projected point, is_inside and triangle
index for every point (TriMesh::project_local_point, point_composite_shape.rs:59-80);
feature, time of impact and normal for every culled ray
(TriMesh::cast_local_ray_with_culling, ray_trimesh.rs:152-178).
"""
import numpy as np
import pytest

from parry_amd import TriMesh, scenes, split_hits, split_projections

pytestmark = pytest.mark.gpu


def _project(mesh, pts, solid):
    import torch

    out = mesh.project_local_points(torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda(), solid)
    torch.cuda.synchronize()
    p, ins, tri = split_projections(out)
    return p.cpu().numpy(), ins.cpu().numpy().view(np.uint32), tri.cpu().numpy().view(np.uint32)


def _check_projection(label, got, ref):
    p, ins, tri = got
    assert np.array_equal(tri, ref["triangle"]), f"{label}: {np.count_nonzero(tri != ref['triangle'])} triangles differ"
    assert np.array_equal(p, ref["point"]), f"{label}: projected points differ"
    assert np.array_equal(ins, ref["is_inside"]), f"{label}: is_inside differs"


def test_projection_matches_oracle(orc):
    v, f = scenes.noisy_icosphere(5, seed=1)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    rng = np.random.default_rng(7)
    centroids = v[f].mean(axis=1)
    pts = np.concatenate([rng.uniform(-1.6, 1.6, (200000, 3)), rng.normal(0.0, 0.4, (50000, 3)),
                          centroids, v]).astype(np.float32)
    for solid in (False, True):
        print(f"case: projection, icosphere s=5, {pts.shape[0]} points, solid={solid}", flush=True)
        ref = om.project_points(pts, solid)
        _check_projection(f"solid={solid}", _project(mesh, pts, solid), ref)
        print(f"  {int(ref['is_inside'].sum())} inside, W/point {ref['W'] / len(pts):.2f}, "
              f"K/point {ref['K'] / len(pts):.2f}")


def test_projection_edge_cases(orc):
    import torch

    # one triangle: the Bvh's partial root (bvh_traverse.rs:348-359)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    pts = np.array([[0.25, 0.5, 3], [0.25, 0.5, 0], [-1, -2, 0.5], [3, -0.5, 1], [1, 1, 0], [1.4, 1.0, 0]], np.float32)
    mesh, om = TriMesh(tri, idx), orc.OracleTriMesh(tri, idx)
    print("case: projection, single triangle", flush=True)
    for solid in (False, True):
        _check_projection("single", _project(mesh, pts, solid), om.project_points(pts, solid))
    # degenerate (collinear) triangles next to a proper one: solid changes the answer
    v = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.5, 2.5, (4096, 3)).astype(np.float32)
    pts[:64, 1:] = 0.0  # on the degenerate triangle's line
    print("case: projection, degenerate triangles", flush=True)
    for solid in (False, True):
        _check_projection(f"degenerate solid={solid}", _project(mesh, pts, solid), om.project_points(pts, solid))
    # no points
    out = mesh.project_local_points(torch.zeros((0, 3), dtype=torch.float32, device="cuda"), False)
    torch.cuda.synchronize()
    assert out.shape == (0, 5)


def _cast_culled(mesh, rays, mode):
    import torch

    out = mesh.cast_local_rays_with_culling(torch.from_numpy(rays).cuda(), 10.0, mode)
    torch.cuda.synchronize()
    toi, nrm, feat = split_hits(out)
    return toi.cpu().numpy(), nrm.cpu().numpy(), feat.cpu().numpy().view(np.uint32)


def test_culling_matches_oracle(orc):
    v, f = scenes.noisy_icosphere(5, seed=1)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    outside = scenes.rays_into_unit_ball(200000, seed=8)
    inside = scenes.rays_into_unit_ball(60000, seed=9, origin_radius=0.5)  # start inside: back faces first
    rays = np.concatenate([outside, inside]).astype(np.float32)
    for code, mode in ((1, "ignore_backfaces"), (2, "ignore_frontfaces")):
        ref = om.cast_rays_with_culling(rays, 10.0, code)
        for mapping in ("static", "binned"):  # the mesh's binned mapping sorts culled casts too
            print(f"case: culling {mode}, {rays.shape[0]} rays, {mapping}", flush=True)
            mesh.set_tuning(mapping)
            toi, nrm, feat = _cast_culled(mesh, rays, mode)
            assert np.array_equal(feat, ref["feature"]), f"{mode}: {np.count_nonzero(feat != ref['feature'])} features differ"
            hit = feat != 0xFFFFFFFF
            assert np.array_equal(toi[hit], ref["toi"][hit]) and np.array_equal(nrm[hit], ref["normal"][hit])
            print(f"  {int(hit.sum())} hits")


def test_culling_reference_kat():
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    mesh = TriMesh(tri, np.array([[0, 1, 2]], np.uint32))
    rays = np.array([[0, 0, -1, 0, 0, 1], [0, 0, 1, 0, 0, -1]], np.float32)  # ray_up, ray_down (ray_trimesh.rs:187-212)
    miss = 0xFFFFFFFF
    _, _, front = _cast_culled(mesh, rays, "ignore_frontfaces")
    _, _, back = _cast_culled(mesh, rays, "ignore_backfaces")
    assert front[0] != miss and front[1] == miss
    assert back[0] == miss and back[1] != miss


def test_traversal_counts_match_oracle(orc):
    """The instrumented launches fetch exactly the oracle's wide records and test
    exactly its triangles (summed over the queries)."""
    import torch

    v, f = scenes.noisy_icosphere(5, seed=1)
    mesh, om = TriMesh(v, f), orc.OracleTriMesh(v, f)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1.6, 1.6, (65536, 3)).astype(np.float32)
    counts = torch.zeros((pts.shape[0], 2), dtype=torch.int32, device="cuda")
    print("case: projection counts, 65536 points", flush=True)
    mesh.project_local_points(torch.from_numpy(pts).cuda(), False, counts=counts)
    torch.cuda.synchronize()
    c = counts.cpu().numpy().astype(np.int64).sum(axis=0)
    ref = om.project_points(pts, False)
    assert (int(c[0]), int(c[1])) == (ref["W"], ref["K"])
    rays = scenes.rays_into_unit_ball(65536, seed=12)
    ref = om.cast_rays_with_culling(rays, 10.0, 2)
    for mapping in ("static", "binned"):
        counts = torch.zeros((rays.shape[0], 2), dtype=torch.int32, device="cuda")
        print(f"case: culling counts, 65536 rays, {mapping}", flush=True)
        mesh.set_tuning(mapping)
        mesh.cast_local_rays_with_culling(torch.from_numpy(rays).cuda(), 10.0, "ignore_frontfaces", counts=counts)
        torch.cuda.synchronize()
        c = counts.cpu().numpy().astype(np.int64).sum(axis=0)
        assert (int(c[0]), int(c[1])) == (ref["W"], ref["K"])
