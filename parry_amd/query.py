"""Query types mirroring parry3d ``src/query`` (ray casts; contacts are added beside them)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

Vec3 = Tuple[float, float, float]

# parry3d::query::Ray, #[repr(C)] f32 (src/query/ray/ray.rs:75-88).
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3)])
# pq_ray_hit (include/parry_mi355x.h): RayIntersection with the TriMesh feature index.
RAY_HIT_DTYPE = np.dtype([("time_of_impact", "<f4"), ("normal", "<f4", 3), ("feature", "<u4")])
assert RAY_DTYPE.itemsize == 24 and RAY_HIT_DTYPE.itemsize == 20


@dataclass(frozen=True)
class Ray:
    """parry3d ``Ray`` (ray.rs:67-88): origin and (not necessarily unit) direction."""

    origin: Vec3
    dir: Vec3

    def point_at(self, t: float) -> Vec3:
        return tuple(o + d * t for o, d in zip(self.origin, self.dir))


@dataclass(frozen=True)
class FeatureId:
    """parry3d ``FeatureId`` (src/shape/feature_id.rs:117): kind in {Vertex, Edge, Face, Unknown}."""

    kind: str
    index: int = 0

    @staticmethod
    def face(i: int) -> "FeatureId":
        return FeatureId("Face", i)


@dataclass(frozen=True)
class RayIntersection:
    """parry3d ``RayIntersection`` (ray.rs:286-315)."""

    time_of_impact: float
    normal: Vec3
    feature: FeatureId


# ---------------------------------------------------------------------------
# Contacts
# ---------------------------------------------------------------------------

# pq_contact (include/parry_mi355x.h): parry's Contact in world space + has_contact.
CONTACT_DTYPE = np.dtype(
    [("point1", "<f4", 3), ("point2", "<f4", 3), ("normal1", "<f4", 3), ("normal2", "<f4", 3), ("dist", "<f4"),
     ("has_contact", "<u4")]
)
assert CONTACT_DTYPE.itemsize == 56
CONTACT_UNSUPPORTED = 0xFFFFFFFF


@dataclass(frozen=True)
class Contact:
    """parry3d ``Contact`` (src/query/contact/contact.rs:71-105)."""

    point1: Vec3
    point2: Vec3
    normal1: Vec3
    normal2: Vec3
    dist: float


def _check_pairs(fn: str, pairs, num_shapes: int, check_indices: bool) -> None:
    """(m, 2) int32/uint32 pairs; with check_indices, every index below num_shapes
    (one device reduction and a synchronisation)."""
    import torch

    if pairs.dtype not in (torch.int32, getattr(torch, "uint32", torch.int32)) or pairs.dim() != 2 \
            or pairs.shape[1] != 2:
        raise ValueError(f"{fn}: pairs must be an (m, 2) int32/uint32 tensor, got {tuple(pairs.shape)} {pairs.dtype}")
    if check_indices and pairs.numel():
        v = pairs.view(torch.int32)
        lo, hi = int(v.min()), int(v.max())
        if lo < 0 or hi >= num_shapes:
            raise ValueError(f"{fn}: pair indices must lie in [0, {num_shapes}), got [{lo}, {hi}]")


def _check_out(fn: str, t, shape: tuple, dtype) -> None:
    if not (t.is_cuda and t.is_contiguous()) or tuple(t.shape) != shape or t.dtype != dtype:
        raise ValueError(f"{fn}: out must be a contiguous CUDA {dtype} tensor of shape {shape}")


def _check_workspace(fn: str, ws, need: int) -> None:
    if ws is not None and (not ws.is_cuda or not ws.is_contiguous() or ws.numel() * ws.element_size() < need):
        raise ValueError(f"{fn}: workspace must be a contiguous CUDA tensor of at least {need} bytes")


def contact_pairs(shapes, poses, pairs, prediction: float, init_dirs=None, with_iters: bool = False, out=None,
                  workspace=None, stream=None, check_indices: bool = True):
    """Batched ``query::contact`` (contact_shape_shape.rs:123-138) on CUDA tensors.

    ``shapes``: a ``ShapeSet``; ``poses``: (n, 7) f32 (quaternion i j k w,
    translation); ``pairs``: (m, 2) int32/uint32; ``init_dirs``: optional (m, 3)
    f32 GJK seeds in shape 1's local frame.  Returns a (m, 14) f32 tensor of
    ``pq_contact`` records (view ``has_contact`` with ``.view(torch.int32)[:, 13]``)
    and, if ``with_iters``, a (m, 2) int32 tensor of GJK / EPA iterations.
    """
    import ctypes

    import torch

    from . import _native

    m = pairs.shape[0]
    dev = poses.device
    if out is None:
        out = torch.empty((m, 14), dtype=torch.float32, device=dev)
    iters = torch.empty((m, 2), dtype=torch.int32, device=dev) if with_iters else None
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    for t in (poses, pairs) + ((init_dirs,) if init_dirs is not None else ()):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("contact_pairs: tensors must be contiguous CUDA tensors")
    _check_pairs("contact_pairs", pairs, len(shapes), check_indices)
    _check_out("contact_pairs", out, (m, 14), torch.float32)
    _check_workspace("contact_pairs", workspace, contact_workspace_bytes(m))
    if init_dirs is not None and (tuple(init_dirs.shape) != (m, 3) or init_dirs.dtype != torch.float32):
        raise ValueError("contact_pairs: init_dirs must be float32 (m, 3)")
    _native.check(
        _native.lib().pq_contact_pairs(
            shapes.handle, poses.data_ptr(), pairs.data_ptr(), m, float(prediction),
            None if init_dirs is None else init_dirs.data_ptr(), out.data_ptr(),
            None if iters is None else iters.data_ptr(), None if workspace is None else workspace.data_ptr(),
            ctypes.c_void_p(s.cuda_stream),
        )
    )
    return (out, iters) if with_iters else out


def contact_workspace_bytes(num_pairs: int) -> int:
    from . import _native

    return int(_native.lib().pq_contact_workspace_bytes(num_pairs))


PQ_CONTACT_TRUNCATED = 0xFFFFFFFE


class MeshContact:
    """``query::contact`` between one ``TriMesh`` and every shape of a ``ShapeSet``
    (contact_composite_shape_shape.rs:14-59, through DefaultQueryDispatcher::contact,
    default_query_dispatcher.rs:343-346): per shape, the first contact with the
    smallest dist among the mesh triangles its loosened AABB selects.
    """

    def __init__(self, mesh, shapes):
        import ctypes

        from . import _native

        self.mesh, self.shapes = mesh, shapes  # kept alive: the handle points into both
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_mesh_contact_create(mesh._h, shapes.handle, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                from . import _native

                _native.lib().pq_mesh_contact_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def workspace_bytes(self, max_candidates: int) -> int:
        from . import _native

        return int(_native.lib().pq_mesh_contact_workspace_bytes(self._h, int(max_candidates)))

    def contact(self, mesh_pose, poses, prediction: float, max_candidates: int, out=None, triangles=None,
                workspace=None, stream=None, needed=None):
        """``mesh_pose``: (7,) and ``poses``: (n, 7) float32 CUDA tensors (i j k w tx ty tz).
        Returns (contacts (n, 14) f32 pq_contact records, triangles (n,) int32, the number of
        candidates the call needed as a (1,) int64 tensor); all asynchronous on ``stream``.
        """
        import ctypes

        import torch

        from . import _native

        n = poses.shape[0]
        dev = poses.device
        for t in (mesh_pose, poses):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise ValueError("mesh_pose and poses must be contiguous CUDA float32 tensors")
        if tuple(mesh_pose.shape) != (7,) or tuple(poses.shape) != (len(self.shapes), 7):
            raise ValueError(f"mesh_pose must be (7,) and poses ({len(self.shapes)}, 7)")
        if int(max_candidates) < 0:
            raise ValueError("max_candidates must be non-negative")
        if out is None:
            out = torch.empty((n, 14), dtype=torch.float32, device=dev)
        if triangles is None:
            triangles = torch.empty(n, dtype=torch.int32, device=dev)
        if workspace is None:
            workspace = torch.empty(self.workspace_bytes(max_candidates), dtype=torch.uint8, device=dev)
        if needed is None:
            needed = torch.zeros(1, dtype=torch.int64, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        _native.check(_native.lib().pq_mesh_contact_shapes(
            self._h, mesh_pose.data_ptr(), poses.data_ptr(), float(prediction), int(max_candidates), out.data_ptr(),
            triangles.data_ptr(), needed.data_ptr(), workspace.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        return out, triangles, needed


# ShapeCastStatus (shape_cast.rs:12-33) numbered from 1; PQ_CAST_NONE is None.
PQ_CAST_NONE, PQ_CAST_OUT_OF_ITERATIONS, PQ_CAST_CONVERGED, PQ_CAST_FAILED, PQ_CAST_PENETRATING = range(5)
F32_MAX = 3.4028234663852886e38


def cast_shapes_workspace_bytes(num_pairs: int) -> int:
    from . import _native

    return int(_native.lib().pq_cast_shapes_workspace_bytes(int(num_pairs)))


def cast_shapes_pairs(shapes, poses, vels, pairs, max_time_of_impact: float = F32_MAX, target_distance: float = 0.0,
                      stop_at_penetration: bool = True, compute_impact_geometry_on_penetration: bool = True,
                      with_iters: bool = False, out=None, workspace=None, stream=None, check_indices: bool = True):
    """Batched ``query::cast_shapes`` (shape_cast.rs:274-286) on CUDA tensors, with
    ``ShapeCastOptions``' fields as keyword arguments (parry's defaults).

    ``shapes``: a ``ShapeSet``; ``poses``: (n, 7) f32; ``vels``: (n, 3) f32 linear
    velocities; ``pairs``: (m, 2) int32/uint32.  Returns a (m, 14) f32 tensor of
    ``pq_shape_cast_hit`` records (toi, witness1, witness2, normal1, normal2 in the
    shapes' local frames; view the status with ``.view(torch.int32)[:, 13]``) and,
    if ``with_iters``, the (m,) int32 GJK ray-cast iterations.
    """
    import ctypes

    import torch

    from . import _native

    m = pairs.shape[0]
    dev = poses.device
    for t in (poses, vels, pairs):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("cast_shapes_pairs: tensors must be contiguous CUDA tensors")
    if poses.dtype != torch.float32 or vels.dtype != torch.float32:
        raise ValueError("cast_shapes_pairs: poses and vels must be float32")
    if tuple(poses.shape) != (len(shapes), 7) or tuple(vels.shape) != (len(shapes), 3):
        raise ValueError(f"cast_shapes_pairs: poses must be ({len(shapes)}, 7) and vels ({len(shapes)}, 3)")
    _check_pairs("cast_shapes_pairs", pairs, len(shapes), check_indices)
    _check_workspace("cast_shapes_pairs", workspace, cast_shapes_workspace_bytes(m))
    if out is None:
        out = torch.empty((m, 14), dtype=torch.float32, device=dev)
    iters = torch.empty(m, dtype=torch.int32, device=dev) if with_iters else None
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    opts = _native.ShapeCastOptions(float(max_time_of_impact), float(target_distance), int(bool(stop_at_penetration)),
                                    int(bool(compute_impact_geometry_on_penetration)))
    _native.check(_native.lib().pq_cast_shapes_pairs(
        shapes.handle, poses.data_ptr(), vels.data_ptr(), pairs.data_ptr(), m, ctypes.byref(opts), out.data_ptr(),
        None if iters is None else iters.data_ptr(), None if workspace is None else workspace.data_ptr(),
        ctypes.c_void_p(s.cuda_stream)))
    return (out, iters) if with_iters else out


# pq_contact_manifold: local_n1, local_n2, num_points, status, then 25 pq_tracked_contact.
MANIFOLD_MAX_POINTS = 25
TRACKED_CONTACT_DTYPE = np.dtype([("local_p1", "<f4", 3), ("local_p2", "<f4", 3), ("dist", "<f4"),
                                  ("fid1", "<u4"), ("fid2", "<u4")])
CONTACT_MANIFOLD_DTYPE = np.dtype([("local_n1", "<f4", 3), ("local_n2", "<f4", 3), ("num_points", "<u4"),
                                   ("status", "<u4"), ("points", TRACKED_CONTACT_DTYPE, MANIFOLD_MAX_POINTS)])
assert CONTACT_MANIFOLD_DTYPE.itemsize == 932


def contact_manifolds_workspace_bytes(num_pairs: int) -> int:
    from . import _native

    return int(_native.lib().pq_contact_manifolds_workspace_bytes(int(num_pairs)))


def contact_manifolds_pfm(shapes, poses, pairs, prediction: float, out=None, workspace=None, stream=None,
                          check_indices: bool = True):
    """Batched ``contact_manifold_pfm_pfm`` (contact_manifolds_pfm_pfm.rs:44-162) on fresh
    manifolds, CUDA tensors in and out.  Returns a uint8 CUDA tensor of (m, 932)
    ``pq_contact_manifold`` records; view them with ``CONTACT_MANIFOLD_DTYPE`` after a
    copy to the host."""
    import ctypes

    import torch

    from . import _native

    m = pairs.shape[0]
    for t in (poses, pairs):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("contact_manifolds_pfm: tensors must be contiguous CUDA tensors")
    if tuple(poses.shape) != (len(shapes), 7) or poses.dtype != torch.float32:
        raise ValueError(f"contact_manifolds_pfm: poses must be float32 ({len(shapes)}, 7)")
    _check_pairs("contact_manifolds_pfm", pairs, len(shapes), check_indices)
    _check_workspace("contact_manifolds_pfm", workspace, contact_manifolds_workspace_bytes(m))
    if out is None:
        out = torch.empty((m, CONTACT_MANIFOLD_DTYPE.itemsize), dtype=torch.uint8, device=poses.device)
    _check_out("contact_manifolds_pfm", out, (m, CONTACT_MANIFOLD_DTYPE.itemsize), torch.uint8)
    s = stream if stream is not None else torch.cuda.current_stream(poses.device)
    _native.check(_native.lib().pq_contact_manifolds_pfm_pairs(
        shapes.handle, poses.data_ptr(), pairs.data_ptr(), m, float(prediction), out.data_ptr(),
        None if workspace is None else workspace.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
    return out


MANIFOLD_UPDATED = 2  # PQ_MANIFOLD_UPDATED: try_update_contacts_eps kept the manifold


def contact_manifolds_pfm_update(shapes, poses, pairs, prediction: float, manifolds, workspace=None, stream=None,
                                 check_indices: bool = True):
    """``contact_manifold_pfm_pfm`` (contact_manifolds_pfm_pfm.rs:44-162) on last frame's
    manifolds, in place: ``manifolds`` is the (m, 932) uint8 CUDA tensor of
    ``pq_contact_manifold`` records a previous call left (zeros for new pairs).  Kept
    manifolds (try_update_contacts_eps, contact_manifold.rs:662-699) come back with
    status ``MANIFOLD_UPDATED``; the others are recomputed with GJK seeded by the
    cached normal.  Returns ``manifolds``."""
    import ctypes

    import torch

    from . import _native

    m = pairs.shape[0]
    for t in (poses, pairs, manifolds):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("contact_manifolds_pfm_update: tensors must be contiguous CUDA tensors")
    if tuple(poses.shape) != (len(shapes), 7) or poses.dtype != torch.float32:
        raise ValueError(f"contact_manifolds_pfm_update: poses must be float32 ({len(shapes)}, 7)")
    if tuple(manifolds.shape) != (m, CONTACT_MANIFOLD_DTYPE.itemsize) or manifolds.dtype != torch.uint8:
        raise ValueError(f"contact_manifolds_pfm_update: manifolds must be uint8 ({m}, {CONTACT_MANIFOLD_DTYPE.itemsize})")
    _check_pairs("contact_manifolds_pfm_update", pairs, len(shapes), check_indices)
    _check_workspace("contact_manifolds_pfm_update", workspace, contact_manifolds_workspace_bytes(m))
    s = stream if stream is not None else torch.cuda.current_stream(poses.device)
    _native.check(_native.lib().pq_contact_manifolds_pfm_update(
        shapes.handle, poses.data_ptr(), pairs.data_ptr(), m, float(prediction), manifolds.data_ptr(),
        None if workspace is None else workspace.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
    return manifolds
