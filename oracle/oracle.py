"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  ``ensure_built()`` runs the oracle's Makefile
when the shared object is missing.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_int = ctypes.c_int

_SIG = {
    "orc_trimesh_create_f32": (_vp, [_vp, _u32, _vp, _u32]),
    "orc_trimesh_create_f64": (_vp, [_vp, _u32, _vp, _u32]),
    "orc_trimesh_free_f32": (None, [_vp]),
    "orc_trimesh_free_f64": (None, [_vp]),
    "orc_trimesh_num_nodes_f32": (_u32, [_vp]),
    "orc_trimesh_depth_f32": (_u32, [_vp]),
    "orc_trimesh_nodes_f32": (None, [_vp, _vp]),
    "orc_cast_rays_f32": (None, [_vp, _vp, _u64, _f32, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int]),
    "orc_cast_rays_f64": (None, [_vp, _vp, _u64, _f64, _int, _vp, _vp, _vp, _vp, _int]),
    "orc_cast_rays_brute_f32": (None, [_vp, _vp, _u64, _f32, _vp, _vp, _vp, _int]),
    "orc_aabb_cast_ray_f32": (_int, [_vp, _vp, _vp, _f32, _int, _vp]),
    "orc_triangle_cast_ray_f32": (_int, [_vp, _vp, _f32, _vp, _vp]),
    "orc_cast_rays_culling_f32": (None, [_vp, _vp, _u64, _f32, _int, _vp, _vp, _vp, _vp, _int, _int]),
    "orc_cast_rays_culling_f64": (None, [_vp, _vp, _u64, _f64, _int, _vp, _vp, _vp, _vp, _int]),
    "orc_project_points_f32": (None, [_vp, _vp, _u64, _int, _vp, _vp, _vp, _vp, _int, _int]),
    "orc_project_points_f64": (None, [_vp, _vp, _u64, _int, _vp, _vp, _vp, _vp, _int, _int]),
    "orc_triangle_project_point_f32": (_int, [_vp, _vp, _int, _vp]),
    "orc_trimesh_contact_shapes_f32": (None, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _f32, _vp, _vp, _vp, _vp, _int]),
    "orc_trimesh_contact_shapes_f64": (None, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _f64, _vp, _vp, _vp, _vp, _int]),
    "orc_manifolds_pfm_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp, _int]),
    "orc_manifolds_pfm_f64": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f64, _vp, _vp, _vp, _vp, _vp, _int]),
    "orc_manifolds_pfm_update_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp,
                                            _int]),
    "orc_manifolds_pfm_faces_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp, _int,
                                           _vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "orc_manifolds_pfm_faces_update_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp, _int,
                                           _vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "orc_bvh_create_ploc_f32": (_vp, [_vp, _u32]),
    "orc_ploc_keys_f32": (None, [_vp, _u32, _vp]),
    "orc_morton_encode_unorm": (_u64, [_f64, _f64, _f64]),
    "orc_cast_shapes_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, _u64, _f32, _f32, _int, _int, _vp, _vp, _vp,
                                   _int]),
    "orc_cast_shapes_f64": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, _u64, _f64, _f64, _int, _int, _vp, _vp, _vp,
                                   _int]),
}

_lib = None


def ensure_built() -> str:
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        ensure_built()
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


class OracleTriMesh:
    """parry3d ``TriMesh`` restated on the CPU, in f32 (parry3d) or f64 (parry3d-f64)."""

    def __init__(self, vertices: np.ndarray, indices: np.ndarray, f64: bool = False):
        self.f64 = f64
        dt = np.float64 if f64 else np.float32
        self.v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
        self.i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        l = lib()
        create = l.orc_trimesh_create_f64 if f64 else l.orc_trimesh_create_f32
        self.h = create(self.v.ctypes.data, self.v.shape[0], self.i.ctypes.data, self.i.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            (lib().orc_trimesh_free_f64 if self.f64 else lib().orc_trimesh_free_f32)(self.h)
            self.h = None

    def num_triangles(self) -> int:
        return self.i.shape[0]

    def bvh_nodes(self) -> np.ndarray:
        """Depth-first BvhNodeWide records (64-byte layout), f32 only."""
        from parry_amd.shape import BVH_NODE_WIDE_DTYPE

        assert not self.f64
        n = lib().orc_trimesh_num_nodes_f32(self.h)
        out = np.zeros(n, dtype=BVH_NODE_WIDE_DTYPE)
        lib().orc_trimesh_nodes_f32(self.h, out.ctypes.data)
        return out

    def depth(self) -> int:
        assert not self.f64
        return lib().orc_trimesh_depth_f32(self.h)

    def cast_rays(self, rays: np.ndarray, max_toi: float, solid: bool = True, threads: int | None = None,
                  per_ray_counts: bool = False, brute: bool = False):
        """Returns dict(feature u32 [n], toi [n], normal [n,3], W, K, max_stack[, w, k])."""
        dt = np.float64 if self.f64 else np.float32
        r = np.ascontiguousarray(rays, dtype=dt).reshape(-1, 6)
        n = r.shape[0]
        fid = np.zeros(n, np.uint32)
        toi = np.zeros(n, dt)
        nrm = np.zeros((n, 3), dt)
        cnt = np.zeros(3, np.uint64)
        th = default_threads() if threads is None else threads
        l = lib()
        out = {"feature": fid, "toi": toi, "normal": nrm}
        if brute:
            assert not self.f64
            l.orc_cast_rays_brute_f32(self.h, r.ctypes.data, n, max_toi, fid.ctypes.data, toi.ctypes.data, nrm.ctypes.data, th)
            return out
        if self.f64:
            l.orc_cast_rays_f64(self.h, r.ctypes.data, n, max_toi, int(solid), fid.ctypes.data, toi.ctypes.data,
                                nrm.ctypes.data, cnt.ctypes.data, th)
        else:
            w = np.zeros(n, np.uint32) if per_ray_counts else None
            k = np.zeros(n, np.uint32) if per_ray_counts else None
            l.orc_cast_rays_f32(self.h, r.ctypes.data, n, max_toi, int(solid), fid.ctypes.data, toi.ctypes.data,
                                nrm.ctypes.data, cnt.ctypes.data, None if w is None else w.ctypes.data,
                                None if k is None else k.ctypes.data, th)
            if per_ray_counts:
                out["w"], out["k"] = w, k
        out["W"], out["K"], out["max_stack"] = int(cnt[0]), int(cnt[1]), int(cnt[2])
        return out

    def cast_rays_with_culling(self, rays: np.ndarray, max_toi: float, culling: int, threads: int | None = None,
                               brute: bool = False):
        """TriMesh::cast_local_ray_with_culling (ray_trimesh.rs:152-178): culling 1 =
        IgnoreBackfaces, 2 = IgnoreFrontfaces.  Returns dict(feature, toi, normal, W, K): W and K are
        the Bvh wide records fetched and the triangles tested, summed over the rays."""
        assert culling in (1, 2) and not (brute and self.f64)
        dt = np.float64 if self.f64 else np.float32
        r = np.ascontiguousarray(rays, dtype=dt).reshape(-1, 6)
        n = r.shape[0]
        fid = np.zeros(n, np.uint32)
        toi = np.zeros(n, dt)
        nrm = np.zeros((n, 3), dt)
        cnt = np.zeros(3, np.uint64)
        th = default_threads() if threads is None else threads
        if self.f64:
            lib().orc_cast_rays_culling_f64(self.h, r.ctypes.data, n, max_toi, int(culling), fid.ctypes.data,
                                            toi.ctypes.data, nrm.ctypes.data, cnt.ctypes.data, th)
        else:
            lib().orc_cast_rays_culling_f32(self.h, r.ctypes.data, n, max_toi, int(culling), fid.ctypes.data,
                                            toi.ctypes.data, nrm.ctypes.data, cnt.ctypes.data, th, int(brute))
        return {"feature": fid, "toi": toi, "normal": nrm, "W": int(cnt[0]), "K": int(cnt[1])}

    def project_points(self, points: np.ndarray, solid: bool, threads: int | None = None, brute: bool = False):
        """TriMesh::project_local_point (point_composite_shape.rs:59-80, 164-168).
        Returns dict(point [n,3], is_inside u32 [n], triangle u32 [n], W, K): W and K are the Bvh wide
        records fetched and the triangles tested, summed over the points (0 with brute)."""
        dt = np.float64 if self.f64 else np.float32
        p = np.ascontiguousarray(points, dtype=dt).reshape(-1, 3)
        n = p.shape[0]
        proj = np.zeros((n, 3), dt)
        inside = np.zeros(n, np.uint32)
        tri = np.zeros(n, np.uint32)
        cnt = np.zeros(2, np.uint64)
        th = default_threads() if threads is None else threads
        f = lib().orc_project_points_f64 if self.f64 else lib().orc_project_points_f32
        f(self.h, p.ctypes.data, n, int(solid), proj.ctypes.data, inside.ctypes.data, tri.ctypes.data,
          cnt.ctypes.data, th, int(brute))
        return {"point": proj, "is_inside": inside, "triangle": tri, "W": int(cnt[0]), "K": int(cnt[1])}


def _trimesh_contact_shapes(self, mesh_pose, descs, vertices, poses, prediction: float, threads: int | None = None):
    """query::contact(mesh_pose, this TriMesh, poses[q], shape q) for every shape q
    (contact_composite_shape_shape.rs:14-59).  Returns dict(status int32 [n] (1 contact,
    0 None), out [n,13] (point1, point2, normal1, normal2, dist in world space),
    triangle u32 [n], candidates: triangles tested in total)."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    dt = np.float64 if self.f64 else np.float32
    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=dt).reshape(-1, 7)
    mp = np.ascontiguousarray(mesh_pose, dtype=dt).reshape(7)
    n = d.shape[0]
    out = np.zeros((n, 13), dt)
    status = np.zeros(n, np.int32)
    tri = np.zeros(n, np.uint32)
    cand = np.zeros(1, np.uint64)
    th = default_threads() if threads is None else threads
    f = lib().orc_trimesh_contact_shapes_f64 if self.f64 else lib().orc_trimesh_contact_shapes_f32
    f(self.h, mp.ctypes.data, d.ctypes.data, n, v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data,
      prediction, out.ctypes.data, status.ctypes.data, tri.ctypes.data, cand.ctypes.data, th)
    return {"status": status, "out": out, "triangle": tri, "candidates": int(cand[0])}


OracleTriMesh.contact_shapes = _trimesh_contact_shapes


def triangle_project_point(abc, pt, solid: bool):
    """Triangle::project_local_point (point_triangle.rs:54-291), f32: (point, is_inside)."""
    t = np.ascontiguousarray(abc, np.float32).reshape(9)
    q = np.ascontiguousarray(pt, np.float32).reshape(3)
    out = np.zeros(3, np.float32)
    ins = lib().orc_triangle_project_point_f32(t.ctypes.data, q.ctypes.data, int(solid), out.ctypes.data)
    return out, bool(ins)


def aabb_cast_ray(mins, maxs, ray, max_toi: float, solid: bool):
    a = np.asarray(mins, np.float32)
    b = np.asarray(maxs, np.float32)
    r = np.asarray(ray, np.float32)
    out = np.zeros(1, np.float32)
    ok = lib().orc_aabb_cast_ray_f32(a.ctypes.data, b.ctypes.data, r.ctypes.data, max_toi, int(solid), out.ctypes.data)
    return float(out[0]) if ok else None


def triangle_cast_ray(abc, ray, max_toi: float):
    t = np.asarray(abc, np.float32).reshape(9)
    r = np.asarray(ray, np.float32)
    out = np.zeros(4, np.float32)
    fid = np.zeros(1, np.uint32)
    ok = lib().orc_triangle_cast_ray_f32(t.ctypes.data, r.ctypes.data, max_toi, out.ctypes.data, fid.ctypes.data)
    return (out, int(fid[0])) if ok else None


# ---------------------------------------------------------------------------
# Broadphase
# ---------------------------------------------------------------------------

_SIG.update({
    "orc_polyhedra_aabbs_f32": (None, [_vp, _vp, _u32, _vp, _vp, _int]),
    "orc_polyhedra_aabbs_f64": (None, [_vp, _vp, _u32, _vp, _vp, _int]),
    "orc_bvh_create_f32": (_vp, [_vp, _u32]),
    "orc_bvh_free_f32": (None, [_vp]),
    "orc_bvh_num_nodes_f32": (_u32, [_vp]),
    "orc_bvh_nodes_f32": (None, [_vp, _vp]),
    "orc_bvh_update_refit_f32": (None, [_vp, _vp]),
    "orc_bvh_self_pairs_f32": (None, [_vp, _vp, _u64, _vp, _vp]),
    "orc_brute_pairs_f32": (None, [_vp, _u32, _vp, _u64, _vp]),
})


def polyhedra_aabbs(vertices, offsets, poses, f64: bool = False, threads: int | None = None) -> np.ndarray:
    """ConvexPolyhedron::aabb(pose) for every shape (n, 6)."""
    dt = np.float64 if f64 else np.float32
    v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
    o = np.ascontiguousarray(offsets, dtype=np.uint32)
    p = np.ascontiguousarray(poses, dtype=dt).reshape(-1, 7)
    n = o.shape[0] - 1
    out = np.zeros((n, 6), dtype=dt)
    f = lib().orc_polyhedra_aabbs_f64 if f64 else lib().orc_polyhedra_aabbs_f32
    f(v.ctypes.data, o.ctypes.data, n, p.ctypes.data, out.ctypes.data, default_threads() if threads is None else threads)
    return out


class OracleBvh:
    """parry3d Bvh (f32) restated: from_leaves (Binned, or Ploc with ploc=True),
    update+refit, self-BVTT."""

    def __init__(self, aabbs, ploc: bool = False):
        a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        self.n = a.shape[0]
        create = lib().orc_bvh_create_ploc_f32 if ploc else lib().orc_bvh_create_f32
        self.h = create(a.ctypes.data, self.n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_free_f32(self.h)
            self.h = None

    def nodes(self) -> np.ndarray:
        from parry_amd.shape import BVH_NODE_WIDE_DTYPE

        out = np.zeros(lib().orc_bvh_num_nodes_f32(self.h), dtype=BVH_NODE_WIDE_DTYPE)
        lib().orc_bvh_nodes_f32(self.h, out.ctypes.data)
        return out

    def update_and_refit(self, aabbs) -> None:
        a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        assert a.shape[0] == self.n
        lib().orc_bvh_update_refit_f32(self.h, a.ctypes.data)

    def self_pairs(self):
        """(pairs uint32 (m, 2) in parry's callback order, wide records touched)."""
        cnt = np.zeros(1, np.uint64)
        touched = np.zeros(1, np.uint64)
        lib().orc_bvh_self_pairs_f32(self.h, None, 0, cnt.ctypes.data, touched.ctypes.data)
        out = np.zeros((int(cnt[0]), 2), np.uint32)
        lib().orc_bvh_self_pairs_f32(self.h, out.ctypes.data, out.shape[0], cnt.ctypes.data, touched.ctypes.data)
        return out, int(touched[0])


def brute_pairs(aabbs) -> np.ndarray:
    """Every pair i < j of closed-overlapping AABBs by exhaustive test (small n only)."""
    a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    cnt = np.zeros(1, np.uint64)
    lib().orc_brute_pairs_f32(a.ctypes.data, a.shape[0], None, 0, cnt.ctypes.data)
    out = np.zeros((int(cnt[0]), 2), np.uint32)
    lib().orc_brute_pairs_f32(a.ctypes.data, a.shape[0], out.ctypes.data, out.shape[0], cnt.ctypes.data)
    return out


def pair_set(pairs) -> set:
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    lo, hi = np.minimum(p[:, 0], p[:, 1]), np.maximum(p[:, 0], p[:, 1])
    return set(zip(lo.tolist(), hi.tolist()))


# ---------------------------------------------------------------------------
# Contacts
# ---------------------------------------------------------------------------

_SIG.update({
    "orc_contact_pairs_f32": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp, _int]),
    "orc_contact_pairs_f64": (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f64, _vp, _vp, _vp, _vp, _vp, _int]),
})

SHAPE_TRIANGLE_TEST_ONLY = 5
_SIG["orc_local_support_f32"] = (None, [_vp, _vp, _u32, _vp, _u64, _vp])
_SIG["orc_local_support_f64"] = (None, [_vp, _vp, _u32, _vp, _u64, _vp])


def local_support_points(desc, vertices, dirs, f64: bool = False) -> np.ndarray:
    """SupportMap::local_support_point of one shape (a SHAPE_DESC_DTYPE record) along
    each row of dirs (n, 3)."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    dt = np.float64 if f64 else np.float32
    d = np.ascontiguousarray(np.atleast_1d(desc)[:1], dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices if vertices is not None else np.zeros((0, 3)), dtype=dt).reshape(-1, 3)
    dr = np.ascontiguousarray(dirs, dtype=dt).reshape(-1, 3)
    out = np.zeros_like(dr)
    f = lib().orc_local_support_f64 if f64 else lib().orc_local_support_f32
    f(d.ctypes.data, v.ctypes.data if v.size else None, v.shape[0], dr.ctypes.data, dr.shape[0], out.ctypes.data)
    return out

# EpaExit (parry_oracle_contact.hpp): how EPA::closest_points (epa3.rs:428-651) ended.
EPA_EXIT_NAMES = ("not_run", "converged", "stuck", "numerical", "iter_cap", "heap_empty", "failed", "point")
_SIG["orc_shape_aabbs_f32"] = (None, [_vp, _u32, _vp, _u32, _vp, _vp])


def shape_aabbs(descs, vertices, poses) -> np.ndarray:
    """Shape::compute_aabb(pose) for every shape of any type on the path (n, 6) f32."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=np.float32).reshape(-1, 7)
    out = np.zeros((d.shape[0], 6), np.float32)
    lib().orc_shape_aabbs_f32(d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data,
                              out.ctypes.data)
    return out


def contact_pairs(descs, vertices, poses, pairs, prediction: float, init_dirs=None, f64: bool = False,
                  threads: int | None = None):
    """query::contact on every pair; returns dict(status int32 [m], out [m,13], iters uint32 [m,2],
    epa_exit uint32 [m]).

    out = point1, point2, normal1, normal2 (world frame), dist.  status: 1
    contact, 0 None, -1 pair type not handled by the restatement.  iters: GJK
    iterations, EPA iterations + 1 (0 without EPA).  epa_exit: how EPA ended
    (EPA_EXIT_NAMES), 0 without EPA.
    """
    from parry_amd.shape import SHAPE_DESC_DTYPE

    dt = np.float64 if f64 else np.float32
    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=dt).reshape(-1, 7)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    m = pr.shape[0]
    out = np.zeros((m, 13), dtype=dt)
    status = np.zeros(m, np.int32)
    iters = np.zeros((m, 2), np.uint32)
    epa_exit = np.zeros(m, np.uint32)
    seeds = None
    if init_dirs is not None:
        seeds = np.ascontiguousarray(init_dirs, dtype=dt).reshape(-1, 3)
        assert seeds.shape[0] == m
    f = lib().orc_contact_pairs_f64 if f64 else lib().orc_contact_pairs_f32
    f(d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data, pr.ctypes.data, m,
      prediction, None if seeds is None else seeds.ctypes.data, out.ctypes.data, status.ctypes.data,
      iters.ctypes.data, epa_exit.ctypes.data, default_threads() if threads is None else threads)
    return {"status": status, "out": out, "iters": iters, "epa_exit": epa_exit}


_SIG["orc_contact_flops_f32"] = (None, [_vp, _u32, _vp, _u32, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp])
FLOP_COUNT_NAMES = ("add", "mul", "div", "sqrt", "cmp", "other")


def contact_flops(descs, vertices, poses, pairs, prediction: float, init_dirs=None):
    """The f32 oracle's query::contact on every pair (one thread), counting its
    arithmetic: returns dict(counts {add, mul, div, sqrt, cmp, other}, flops (total
    add + mul + div + sqrt), per_pair uint32 [m], dist float32 [m], status int32 [m])."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=np.float32).reshape(-1, 7)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    m = pr.shape[0]
    seeds = None if init_dirs is None else np.ascontiguousarray(init_dirs, dtype=np.float32).reshape(-1, 3)
    counts = np.zeros(6, np.uint64)
    per_pair = np.zeros(m, np.uint32)
    dist = np.zeros(m, np.float32)
    status = np.zeros(m, np.int32)
    lib().orc_contact_flops_f32(d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0],
                                p.ctypes.data, pr.ctypes.data, m, prediction,
                                None if seeds is None else seeds.ctypes.data, counts.ctypes.data,
                                per_pair.ctypes.data, dist.ctypes.data, status.ctypes.data)
    c = {k: int(x) for k, x in zip(FLOP_COUNT_NAMES, counts)}
    return {"counts": c, "flops": c["add"] + c["mul"] + c["div"] + c["sqrt"], "per_pair": per_pair, "dist": dist,
            "status": status}


CAST_NONE, CAST_OUT_OF_ITERATIONS, CAST_CONVERGED, CAST_FAILED, CAST_PENETRATING = range(5)


def cast_shapes(descs, vertices, poses, vels, pairs, max_toi: float = float("inf"), target_distance: float = 0.0,
                stop_at_penetration: bool = True, compute_impact_geometry_on_penetration: bool = True,
                f64: bool = False, threads: int | None = None):
    """query::cast_shapes (shape_cast.rs:274-286) on every pair with per-shape linear
    velocities; returns dict(status uint32 [m] (0 None, else ShapeCastStatus + 1),
    out [m,13] (toi, witness1, witness2, normal1, normal2 in the shapes' local
    frames), iters uint32 [m] (minkowski_ray_cast passes)).  max_toi = inf means
    parry's default Real::MAX."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    dt = np.float64 if f64 else np.float32
    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=dt).reshape(-1, 7)
    vl = np.ascontiguousarray(vels, dtype=dt).reshape(-1, 3)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    m = pr.shape[0]
    out = np.zeros((m, 13), dt)
    status = np.zeros(m, np.uint32)
    iters = np.zeros(m, np.uint32)
    mt = float(np.finfo(dt).max) if not np.isfinite(max_toi) else max_toi
    f = lib().orc_cast_shapes_f64 if f64 else lib().orc_cast_shapes_f32
    f(d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data, vl.ctypes.data,
      pr.ctypes.data, m, mt, target_distance, int(stop_at_penetration), int(compute_impact_geometry_on_penetration),
      out.ctypes.data, status.ctypes.data, iters.ctypes.data, default_threads() if threads is None else threads)
    return {"status": status, "out": out, "iters": iters}


def ploc_keys(aabbs) -> np.ndarray:
    """The Morton keys the PLOC build sorts the leaves by (bvh_ploc_build.rs:12-21)."""
    a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    out = np.zeros(a.shape[0], np.uint64)
    lib().orc_ploc_keys_f32(a.ctypes.data, a.shape[0], out.ctypes.data)
    return out


def morton_encode_unorm(x: float, y: float, z: float) -> int:
    """morton_encode_u64_unorm (morton.rs:33-40)."""
    return int(lib().orc_morton_encode_unorm(float(x), float(y), float(z)))


MANIFOLD_POINTS = 25


def manifolds_pfm(descs, vertices, poses, pairs, prediction: float, f64: bool = False, threads: int | None = None,
                  faces=None, previous=None):
    """contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162) on a fresh manifold
    for every pair, or, given ``previous`` (the dict an earlier call returned, f32), on
    those manifolds.  Returns dict(status int32 [m] (2 kept by try_update_contacts_eps,
    1 ClosestPoints, 0 NoIntersection, -1 no feature map), n1, n2 [m,3], count uint32
    [m], p1, p2 [m,25,3], dist [m,25], fid1, fid2 uint32 [m,25]) in the shapes' local
    frames."""
    from parry_amd.shape import SHAPE_DESC_DTYPE

    dt = np.float64 if f64 else np.float32
    d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=dt).reshape(-1, 3)
    p = np.ascontiguousarray(poses, dtype=dt).reshape(-1, 7)
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    m = pr.shape[0]
    status = np.zeros(m, np.int32)
    normals = np.zeros((m, 6), dt)
    counts = np.zeros(m, np.uint32)
    pts = np.zeros((m, MANIFOLD_POINTS, 7), dt)
    fids = np.zeros((m, MANIFOLD_POINTS, 2), np.uint32)
    th = default_threads() if threads is None else threads
    if previous is not None:
        assert not f64
        normals[:] = np.concatenate([previous["n1"], previous["n2"]], 1)
        counts[:] = previous["count"]
        pts[:] = np.concatenate([previous["p1"], previous["p2"], previous["dist"][:, :, None]], 2)
        fids[:] = np.stack([previous["fid1"], previous["fid2"]], 2)
    if previous is not None and faces is None:
        lib().orc_manifolds_pfm_update_f32(
            d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data, pr.ctypes.data,
            m, prediction, status.ctypes.data, normals.ctypes.data, counts.ctypes.data, pts.ctypes.data,
            fids.ctypes.data, th)
    elif faces is not None:  # ConvexPolyhedron faces, as scenes.polyhedron_faces returns them (f32 only)
        assert not f64
        fc = faces["faces"]
        nrm = np.ascontiguousarray(fc["normal"], np.float32)
        ff = np.ascontiguousarray(fc["first"], np.uint32)
        fn = np.ascontiguousarray(fc["count"], np.uint32)
        fv = np.ascontiguousarray(faces["face_vertices"], np.uint32)
        fe = np.ascontiguousarray(faces["face_edges"], np.uint32)
        sf = np.ascontiguousarray(faces["shape_faces"], np.uint32)
        fn_faces = lib().orc_manifolds_pfm_faces_update_f32 if previous is not None else lib().orc_manifolds_pfm_faces_f32
        fn_faces(
            d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data, pr.ctypes.data,
            m, prediction, status.ctypes.data, normals.ctypes.data, counts.ctypes.data, pts.ctypes.data,
            fids.ctypes.data, th, nrm.ctypes.data, ff.ctypes.data, fn.ctypes.data, len(fc), fv.ctypes.data,
            fe.ctypes.data, sf.ctypes.data)
    else:
        f = lib().orc_manifolds_pfm_f64 if f64 else lib().orc_manifolds_pfm_f32
        f(d.ctypes.data, d.shape[0], v.ctypes.data if v.size else None, v.shape[0], p.ctypes.data, pr.ctypes.data,
          m, prediction, status.ctypes.data, normals.ctypes.data, counts.ctypes.data, pts.ctypes.data,
          fids.ctypes.data, th)
    return {"status": status, "n1": normals[:, :3], "n2": normals[:, 3:], "count": counts, "p1": pts[:, :, :3],
            "p2": pts[:, :, 3:6], "dist": pts[:, :, 6], "fid1": fids[:, :, 0], "fid2": fids[:, :, 1]}
