"""Shapes backed by device-resident records (parry3d ``src/shape``).

``TriMesh`` mirrors parry3d's ``TriMesh`` (src/shape/trimesh.rs:519-528) for
the ray-cast path: construction builds the same binned Bvh
(trimesh.rs:1159-1171) and uploads it; ``RayCast`` methods
(src/query/ray/ray.rs:359-411) run on the GPU in batches.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native
from .query import Ray, RayIntersection, RAY_DTYPE, RAY_HIT_DTYPE, FeatureId

BVH_NODE_WIDE_DTYPE = np.dtype(
    [
        ("left_mins", "<f4", 3), ("left_children", "<u4"), ("left_maxs", "<f4", 3), ("left_data", "<u4"),
        ("right_mins", "<f4", 3), ("right_children", "<u4"), ("right_maxs", "<f4", 3), ("right_data", "<u4"),
    ]
)
assert BVH_NODE_WIDE_DTYPE.itemsize == 64


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class TriMeshBuilderError(ValueError):
    """parry's TriMeshBuilderError (src/shape/trimesh.rs)."""


class TriMesh:
    """A triangle mesh whose Bvh and triangles live in GPU memory.

    ``vertices``: (nv, 3) float32; ``indices``: (nt, 3) uint32.  ``bvh_nodes``
    optionally supplies parry's Bvh as depth-first ``BVH_NODE_WIDE_DTYPE``
    records (e.g. exported from a Rust ``Bvh``); by default the binned Bvh is
    built here exactly as ``TriMesh::new`` builds it.
    """

    def __init__(self, vertices, indices, bvh_nodes: Optional[np.ndarray] = None):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        if i.shape[0] == 0:
            raise TriMeshBuilderError("EmptyIndices")
        self._vertices = v
        self._indices = i
        lib = _native.lib()
        h = ctypes.c_void_p()
        if bvh_nodes is not None:
            nodes = np.ascontiguousarray(bvh_nodes, dtype=BVH_NODE_WIDE_DTYPE)
            rc = lib.pq_trimesh_create(_ptr(v), v.shape[0], _ptr(i), i.shape[0], _ptr(nodes), nodes.shape[0], ctypes.byref(h))
        else:
            rc = lib.pq_trimesh_create(_ptr(v), v.shape[0], _ptr(i), i.shape[0], None, 0, ctypes.byref(h))
        _native.check(rc)
        self._h = h
        nn, d, nt = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(lib.pq_trimesh_info(h, ctypes.byref(nn), ctypes.byref(d), ctypes.byref(nt)))
        self._num_nodes, self._depth = nn.value, d.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_trimesh_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- accessors -------------------------------------------------------
    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def vertices(self) -> np.ndarray:
        return self._vertices

    def indices(self) -> np.ndarray:
        return self._indices

    def num_triangles(self) -> int:
        return self._indices.shape[0]

    @property
    def bvh_depth(self) -> int:
        return self._depth

    def bvh_nodes(self) -> np.ndarray:
        """The Bvh as depth-first BvhNodeWide records (leaf children = triangle index)."""
        out = np.zeros(self._num_nodes, dtype=BVH_NODE_WIDE_DTYPE)
        _native.check(_native.lib().pq_trimesh_export_bvh(self._h, _ptr(out), out.shape[0]))
        return out

    def to_bytes(self) -> bytes:
        """parry's serialized ``TriMesh`` (serde as bincode 1.x writes it, trimesh.rs:510-528)."""
        from .wire import trimesh_to_bytes

        return trimesh_to_bytes(self._vertices, self._indices, self.bvh_nodes())

    @classmethod
    def from_bytes(cls, data: bytes) -> "TriMesh":
        """The mesh of a serialized parry ``TriMesh``, keeping its Bvh as serialized."""
        from .wire import trimesh_from_bytes

        v, i, nodes = trimesh_from_bytes(data)
        return cls(v, i, bvh_nodes=nodes)

    # -- tuning ----------------------------------------------------------
    def set_tuning(self, mapping: str = "persistent", chunk: int = 0, blocks_per_cu: int = 0, block_size: int = 256,
                   stack_regs: int = 0) -> None:
        """Selects the ray kernel's mapping ("static", "persistent", "wave" or "binned": static lanes
        over the rays sorted by the cell where they enter the Bvh's root box), workgroup size and the
        stack entries kept in registers; results never change."""
        t = _native.ray_tuning(mapping, chunk, blocks_per_cu, block_size, stack_regs)
        _native.check(_native.lib().pq_trimesh_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = _native.RayTuning()
        _native.check(_native.lib().pq_trimesh_get_tuning(self._h, ctypes.byref(t)))
        return _native.ray_tuning_dict(t)

    def autotune(self, rays, max_time_of_impact: float, out=None, stream=None):
        """Times the kernel mappings on these rays and keeps the fastest; returns (tuning, ms per launch)."""
        import torch

        n = rays.shape[0]
        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[1:] == (6,)):
            raise ValueError("rays must be a contiguous CUDA float32 tensor of shape (n, 6)")
        if n == 0:
            raise ValueError("autotune needs at least one ray")
        if out is None:
            out = torch.empty((n, 5), dtype=torch.float32, device=rays.device)
        elif not (out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.numel() >= 5 * n):
            raise ValueError("out must be a contiguous CUDA float32 tensor with room for (n, 5)")
        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        t = _native.RayTuning()
        ms = ctypes.c_float()
        _native.check(_native.lib().pq_trimesh_autotune(
            self._h, rays.data_ptr(), n, float(max_time_of_impact), out.data_ptr(), ctypes.c_void_p(s.cuda_stream),
            ctypes.byref(t), ctypes.byref(ms)))
        return self.tuning(), float(ms.value)

    # -- RayCast ---------------------------------------------------------
    def cast_local_rays_and_get_normal(self, rays, max_time_of_impact: float, solid: bool, out=None, stream=None):
        """Batched ``RayCast::cast_local_ray_and_get_normal``.

        ``rays`` is either a CUDA ``torch.Tensor`` of shape (n, 6) float32 (origin,
        dir) -- then the result is a (n, 5) float32 CUDA tensor of
        ``pq_ray_hit`` records produced asynchronously on ``stream`` (default:
        torch's current stream); see ``split_hits`` -- or a numpy array, for
        which the call copies in and out and returns a ``RAY_HIT_DTYPE``
        structured array.
        """
        lib = _native.lib()
        if isinstance(rays, np.ndarray):
            r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
            hits = np.zeros(r.shape[0], dtype=RAY_HIT_DTYPE)
            _native.check(
                lib.pq_trimesh_cast_local_rays_host(self._h, _ptr(r), r.shape[0], float(max_time_of_impact), int(bool(solid)), _ptr(hits))
            )
            return hits
        import torch

        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[-1] == 6):
            raise ValueError("rays must be a contiguous CUDA float32 tensor of shape (n, 6)")
        n = rays.shape[0]
        if out is None:
            out = torch.empty((n, 5), dtype=torch.float32, device=rays.device)
        elif not (out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.numel() >= 5 * n):
            raise ValueError("out must be a contiguous CUDA float32 tensor with room for (n, 5)")
        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        _native.check(
            lib.pq_trimesh_cast_local_rays(
                self._h, rays.data_ptr(), n, float(max_time_of_impact), int(bool(solid)), out.data_ptr(), ctypes.c_void_p(s.cuda_stream)
            )
        )
        return out

    def cast_local_rays_counted(self, rays, max_time_of_impact: float, solid: bool = True, stream=None):
        """Instrumented batch cast on CUDA tensors: returns (hits (n, 5) f32, counts (n, 2) int32).

        counts[:, 0] = Bvh wide-node records fetched, counts[:, 1] = triangles tested.
        """
        import torch

        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[-1] == 6):
            raise ValueError("rays must be a contiguous CUDA float32 tensor of shape (n, 6)")
        n = rays.shape[0]
        out = torch.empty((n, 5), dtype=torch.float32, device=rays.device)
        counts = torch.empty((n, 2), dtype=torch.int32, device=rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        _native.check(
            _native.lib().pq_trimesh_cast_local_rays_counted(
                self._h, rays.data_ptr(), n, float(max_time_of_impact), int(bool(solid)), out.data_ptr(),
                counts.data_ptr(), ctypes.c_void_p(s.cuda_stream)
            )
        )
        return out, counts

    def cast_local_rays_with_culling(self, rays, max_time_of_impact: float, culling: str, out=None, stream=None,
                                     counts=None):
        """Batched ``TriMesh::cast_local_ray_with_culling`` (ray_trimesh.rs:152-178) on CUDA tensors.

        ``culling`` is "ignore_backfaces" or "ignore_frontfaces" (``RayCullingMode``);
        returns (n, 5) float32 ``pq_ray_hit`` records (see ``split_hits``).  With a
        (n, 2) int32 CUDA tensor ``counts``, the instrumented launch also writes each
        ray's wide-node records fetched and triangles tested.
        """
        import torch

        modes = {"ignore_backfaces": 1, "ignore_frontfaces": 2}
        if culling not in modes:
            raise ValueError(f"culling must be one of {sorted(modes)}")
        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and rays.shape[-1] == 6):
            raise ValueError("rays must be a contiguous CUDA float32 tensor of shape (n, 6)")
        n = rays.shape[0]
        if out is None:
            out = torch.empty((n, 5), dtype=torch.float32, device=rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        lib = _native.lib()
        if counts is None:
            _native.check(lib.pq_trimesh_cast_local_rays_with_culling(
                self._h, rays.data_ptr(), n, float(max_time_of_impact), modes[culling], out.data_ptr(),
                ctypes.c_void_p(s.cuda_stream)))
        else:
            _native.check(lib.pq_trimesh_cast_local_rays_with_culling_counted(
                self._h, rays.data_ptr(), n, float(max_time_of_impact), modes[culling], out.data_ptr(),
                counts.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        return out

    def project_local_points(self, points, solid: bool, out=None, stream=None, counts=None):
        """Batched ``PointQuery::project_local_point`` (point_composite_shape.rs:164-168).

        ``points``: contiguous CUDA float32 tensor (n, 3).  Returns a (n, 5) float32
        tensor of ``pq_point_projection`` records: point xyz, then is_inside and the
        triangle index as uint32 (see ``split_projections``).  With a (n, 2) int32
        CUDA tensor ``counts``, the instrumented launch also writes each point's
        wide-node records fetched and triangles tested.
        """
        import torch

        if not (points.is_cuda and points.dtype == torch.float32 and points.is_contiguous() and points.shape[-1] == 3):
            raise ValueError("points must be a contiguous CUDA float32 tensor of shape (n, 3)")
        n = points.shape[0]
        if out is None:
            out = torch.empty((n, 5), dtype=torch.float32, device=points.device)
        s = stream if stream is not None else torch.cuda.current_stream(points.device)
        lib = _native.lib()
        if counts is None:
            _native.check(lib.pq_trimesh_project_local_points(
                self._h, points.data_ptr(), n, int(bool(solid)), out.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        else:
            _native.check(lib.pq_trimesh_project_local_points_counted(
                self._h, points.data_ptr(), n, int(bool(solid)), out.data_ptr(), counts.data_ptr(),
                ctypes.c_void_p(s.cuda_stream)))
        return out

    def cast_local_ray_and_get_normal(self, ray: Ray, max_time_of_impact: float, solid: bool) -> Optional[RayIntersection]:
        """Single-ray ``RayCast::cast_local_ray_and_get_normal`` (ray.rs:393-403 in local space)."""
        r = np.asarray([list(ray.origin) + list(ray.dir)], dtype=np.float32)
        h = self.cast_local_rays_and_get_normal(r, max_time_of_impact, solid)[0]
        if int(h["feature"]) == _native.PQ_NO_HIT:
            return None
        return RayIntersection(float(h["time_of_impact"]), tuple(float(x) for x in h["normal"]), FeatureId.face(int(h["feature"])))

    def cast_local_ray(self, ray: Ray, max_time_of_impact: float, solid: bool) -> Optional[float]:
        hit = self.cast_local_ray_and_get_normal(ray, max_time_of_impact, solid)
        return None if hit is None else hit.time_of_impact


def split_hits(hits):
    """Views of a (n, 5) float32 ``pq_ray_hit`` tensor: (toi, normal, feature as int32; -1 = None)."""
    import torch

    feat = hits.view(torch.int32)[:, 4]
    return hits[:, 0], hits[:, 1:4], feat


def split_projections(proj):
    """Views of a (n, 5) float32 ``pq_point_projection`` tensor: (point, is_inside, triangle as int32)."""
    import torch

    iv = proj.view(torch.int32)
    return proj[:, 0:3], iv[:, 3], iv[:, 4]


class TriMeshBatch:
    """Many ``TriMesh`` ray-cast batches of uneven sizes in one launch.

    ``ray_counts[i]`` rays are cast against ``meshes[i]``; the rays of all
    meshes are concatenated in mesh order.  Each ray's result is exactly what
    ``meshes[i].cast_local_rays_and_get_normal`` returns for it.  The meshes
    are kept alive by the batch.
    """

    def __init__(self, meshes, ray_counts):
        self._meshes = list(meshes)
        counts = np.asarray(ray_counts, dtype=np.uint64).reshape(-1)
        if counts.shape[0] != len(self._meshes):
            raise ValueError("one ray count per mesh")
        self.offsets = np.zeros(len(self._meshes) + 1, dtype=np.uint64)
        np.cumsum(counts, out=self.offsets[1:])
        handles = (ctypes.c_void_p * len(self._meshes))(*[m.handle.value for m in self._meshes])
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_trimesh_batch_create(handles, len(self._meshes), _ptr(self.offsets), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_trimesh_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def num_rays(self) -> int:
        return int(self.offsets[-1])

    def set_tuning(self, mapping: str = "persistent", chunk: int = 0, blocks_per_cu: int = 0, block_size: int = 256,
                   stack_regs: int = 0) -> None:
        t = _native.ray_tuning(mapping, chunk, blocks_per_cu, block_size, stack_regs)
        _native.check(_native.lib().pq_trimesh_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = _native.RayTuning()
        _native.check(_native.lib().pq_trimesh_batch_get_tuning(self._h, ctypes.byref(t)))
        return _native.ray_tuning_dict(t)

    def _check_rays(self, rays, out):
        import torch

        n = self.num_rays
        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous() and tuple(rays.shape) == (n, 6)):
            raise ValueError(f"rays must be a contiguous CUDA float32 tensor of shape ({n}, 6)")
        if out is None:
            out = torch.empty((n, 5), dtype=torch.float32, device=rays.device)
        elif not (out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.numel() >= 5 * n):
            raise ValueError("out must be a contiguous CUDA float32 tensor with room for (n, 5)")
        return out

    @staticmethod
    def _stream(rays, stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(rays.device)
        return ctypes.c_void_p(s.cuda_stream)

    def cast_local_rays_and_get_normal(self, rays, max_time_of_impact: float, solid: bool, out=None, stream=None):
        """(n, 5) float32 ``pq_ray_hit`` records for the concatenated rays, asynchronously on ``stream``."""
        out = self._check_rays(rays, out)
        _native.check(_native.lib().pq_trimesh_batch_cast_local_rays(
            self._h, rays.data_ptr(), float(max_time_of_impact), int(bool(solid)), out.data_ptr(),
            self._stream(rays, stream)))
        return out

    def cast_local_rays_counted(self, rays, max_time_of_impact: float, solid: bool = True, stream=None):
        import torch

        out = self._check_rays(rays, None)
        counts = torch.empty((self.num_rays, 2), dtype=torch.int32, device=rays.device)
        _native.check(_native.lib().pq_trimesh_batch_cast_local_rays_counted(
            self._h, rays.data_ptr(), float(max_time_of_impact), int(bool(solid)), out.data_ptr(), counts.data_ptr(),
            self._stream(rays, stream)))
        return out, counts

    def autotune(self, rays, max_time_of_impact: float, out=None, stream=None):
        """Times the kernel mappings on these rays and keeps the fastest; returns (tuning, ms per launch)."""
        out = self._check_rays(rays, out)
        t = _native.RayTuning()
        ms = ctypes.c_float()
        _native.check(_native.lib().pq_trimesh_batch_autotune(
            self._h, rays.data_ptr(), float(max_time_of_impact), out.data_ptr(), self._stream(rays, stream),
            ctypes.byref(t), ctypes.byref(ms)))
        return self.tuning(), float(ms.value)

    def capture(self, rays, max_time_of_impact: float, solid: bool, out, stream=None) -> "Graph":
        """Captures one batched cast into a replayable hipGraph writing ``out``."""
        out = self._check_rays(rays, out)
        g = ctypes.c_void_p()
        _native.check(_native.lib().pq_trimesh_batch_capture(
            self._h, rays.data_ptr(), float(max_time_of_impact), int(bool(solid)), out.data_ptr(),
            self._stream(rays, stream), ctypes.byref(g)))
        return Graph(g, keep=(self, rays, out))


class Graph:
    """A captured hipGraph (``pq_graph``); ``launch`` replays it on a stream."""

    def __init__(self, handle: ctypes.c_void_p, keep=()):
        self._h = handle
        self._keep = keep

    def launch(self, stream=None) -> None:
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(_native.lib().pq_graph_launch(self._h, ctypes.c_void_p(s.cuda_stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_graph_destroy(h)
            except Exception:
                pass
            self._h = None


# ---------------------------------------------------------------------------
# Shape sets (ConvexPolyhedron, Ball, ...) for the broadphase and contacts
# ---------------------------------------------------------------------------

SHAPE_BALL = 0
SHAPE_CUBOID = 1
SHAPE_CAPSULE = 2
SHAPE_CONE = 3
SHAPE_CONVEX_POLYHEDRON = 4
SHAPE_TRIANGLE = 5

SHAPE_DESC_DTYPE = np.dtype(
    [("type", "<u4"), ("first_vertex", "<u4"), ("num_vertices", "<u4"), ("reserved", "<u4"), ("params", "<f4", 4)]
)
assert SHAPE_DESC_DTYPE.itemsize == 32


# pq_polyhedron_face: a ConvexPolyhedron face (convex_polyhedron.rs:59-63, 172-185).
POLYHEDRON_FACE_DTYPE = np.dtype([("normal", "<f4", 3), ("first", "<u4"), ("count", "<u4")])
assert POLYHEDRON_FACE_DTYPE.itemsize == 20


class ShapeSet:
    """A device-resident set of shapes (pq_shapes): type tags, parameters and the
    ConvexPolyhedron vertex pool."""

    def __init__(self, descs: np.ndarray, vertices: np.ndarray):
        d = np.ascontiguousarray(descs, dtype=SHAPE_DESC_DTYPE)
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_shapes_create(_ptr(d), d.shape[0], _ptr(v), v.shape[0], ctypes.byref(h)))
        self._h = h
        self.descs = d
        self.vertices = v

    @classmethod
    def convex_polyhedra(cls, vertices: np.ndarray, offsets: np.ndarray) -> "ShapeSet":
        """One ConvexPolyhedron per range offsets[i]..offsets[i+1] of ``vertices``."""
        offs = np.asarray(offsets, dtype=np.uint32)
        n = offs.shape[0] - 1
        d = np.zeros(n, dtype=SHAPE_DESC_DTYPE)
        d["type"] = SHAPE_CONVEX_POLYHEDRON
        d["first_vertex"] = offs[:-1]
        d["num_vertices"] = offs[1:] - offs[:-1]
        return cls(d, vertices)

    def __len__(self) -> int:
        return self.descs.shape[0]

    def set_polyhedron_faces(self, faces: dict) -> None:
        """Gives the ConvexPolyhedra their faces (pq_shapes_set_polyhedron_faces), the
        tables ``ConvexPolyhedron`` keeps (convex_polyhedron.rs:172-185): a dict with
        ``faces`` (POLYHEDRON_FACE_DTYPE), ``face_vertices`` and ``face_edges`` (uint32)
        and ``shape_faces`` (uint32 [n, 2]: first face, face count; 0 faces for other
        shapes), as ``scenes.polyhedron_faces`` builds them."""
        f = np.ascontiguousarray(faces["faces"], dtype=POLYHEDRON_FACE_DTYPE)
        fv = np.ascontiguousarray(faces["face_vertices"], dtype=np.uint32)
        fe = np.ascontiguousarray(faces["face_edges"], dtype=np.uint32)
        sf = np.ascontiguousarray(faces["shape_faces"], dtype=np.uint32).reshape(-1, 2)
        if fv.shape != fe.shape or sf.shape[0] != len(self):
            raise ValueError("set_polyhedron_faces: face_vertices/face_edges differ in length, or shape_faces "
                             f"has {sf.shape[0]} rows for {len(self)} shapes")
        _native.check(_native.lib().pq_shapes_set_polyhedron_faces(
            self._h, _ptr(f), f.shape[0], _ptr(fv), _ptr(fe), fv.shape[0], _ptr(sf)))
        self.faces = {"faces": f, "face_vertices": fv, "face_edges": fe, "shape_faces": sf}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_shapes_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def descriptors(self, out=None, stream=None):
        """(n, 16) f32 CUDA tensor of shape descriptors (the warm-start network's per-shape input)."""
        import torch

        n = len(self)
        if out is None:
            out = torch.empty((n, 16), dtype=torch.float32, device="cuda")
        s = stream if stream is not None else torch.cuda.current_stream(out.device)
        _native.check(_native.lib().pq_shape_descriptors(self._h, out.data_ptr(), ctypes.c_void_p(s.cuda_stream)))
        return out

    def compute_aabbs(self, poses, out=None, stream=None):
        """Shape::compute_aabb for every shape: poses CUDA (n, 7) f32 -> CUDA (n, 6) f32."""
        import torch

        n = len(self)
        if out is None:
            out = torch.empty((n, 6), dtype=torch.float32, device=poses.device)
        s = stream if stream is not None else torch.cuda.current_stream(poses.device)
        _native.check(
            _native.lib().pq_shapes_compute_aabbs(self._h, poses.data_ptr(), out.data_ptr(), ctypes.c_void_p(s.cuda_stream))
        )
        return out
