"""parry's serialized ``Bvh`` and ``TriMesh`` (serde derives written as bincode 1.x
lays them out), through the library's host-only encoders and decoders.

``Bvh`` (bvh_tree.rs:1748-1762): nodes, parents, leaf_node_indices;
``TriMesh`` (trimesh.rs:510-528): bvh, vertices, indices, pseudo_normals,
topology, connected_components, flags -- the last four unset, as
``TriMesh::new`` leaves them.  No GPU is needed.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .shape import BVH_NODE_WIDE_DTYPE


def _nodes(nodes) -> np.ndarray:
    return np.ascontiguousarray(nodes, dtype=BVH_NODE_WIDE_DTYPE).reshape(-1)


def bvh_to_bytes(nodes) -> bytes:
    """Bytes of parry's ``Bvh`` with these depth-first wide nodes."""
    n = _nodes(nodes)
    lib = _native.lib()
    size = ctypes.c_size_t()
    rc = lib.pq_bvh_encode(n.ctypes.data, n.shape[0], None, 0, ctypes.byref(size))
    if rc not in (_native.PQ_OK, _native.PQ_ERR_CAPACITY):
        _native.check(rc)
    buf = ctypes.create_string_buffer(size.value)
    _native.check(lib.pq_bvh_encode(n.ctypes.data, n.shape[0], buf, size.value, ctypes.byref(size)))
    return buf.raw[: size.value]


def bvh_from_bytes(data: bytes) -> np.ndarray:
    """The depth-first wide nodes of a serialized ``Bvh`` (checked against its parents
    and leaf_node_indices)."""
    lib = _native.lib()
    cnt = ctypes.c_uint32()
    rc = lib.pq_bvh_decode(data, len(data), None, 0, ctypes.byref(cnt))
    if rc not in (_native.PQ_OK, _native.PQ_ERR_CAPACITY):
        _native.check(rc)
    out = np.zeros(cnt.value, BVH_NODE_WIDE_DTYPE)
    _native.check(lib.pq_bvh_decode(data, len(data), out.ctypes.data if out.size else None, out.shape[0],
                                    ctypes.byref(cnt)))
    return out


def trimesh_to_bytes(vertices, indices, nodes) -> bytes:
    """Bytes of parry's ``TriMesh`` (as ``TriMesh::new`` makes it) from its parts."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    n = _nodes(nodes)
    lib = _native.lib()
    size = ctypes.c_size_t()
    args = (v.ctypes.data, v.shape[0], i.ctypes.data, i.shape[0], n.ctypes.data, n.shape[0])
    rc = lib.pq_trimesh_encode(*args, None, 0, ctypes.byref(size))
    if rc not in (_native.PQ_OK, _native.PQ_ERR_CAPACITY):
        _native.check(rc)
    buf = ctypes.create_string_buffer(size.value)
    _native.check(lib.pq_trimesh_encode(*args, buf, size.value, ctypes.byref(size)))
    return buf.raw[: size.value]


def trimesh_from_bytes(data: bytes):
    """(vertices float32 (nv, 3), indices uint32 (nt, 3), nodes) of a serialized ``TriMesh``."""
    lib = _native.lib()
    nv, nt, nn = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    rc = lib.pq_trimesh_decode(data, len(data), None, 0, ctypes.byref(nv), None, 0, ctypes.byref(nt), None, 0,
                               ctypes.byref(nn))
    if rc not in (_native.PQ_OK, _native.PQ_ERR_CAPACITY):
        _native.check(rc)
    v = np.zeros((nv.value, 3), np.float32)
    i = np.zeros((nt.value, 3), np.uint32)
    n = np.zeros(nn.value, BVH_NODE_WIDE_DTYPE)
    _native.check(lib.pq_trimesh_decode(data, len(data), v.ctypes.data, v.shape[0], ctypes.byref(nv), i.ctypes.data,
                                        i.shape[0], ctypes.byref(nt), n.ctypes.data, n.shape[0], ctypes.byref(nn)))
    return v, i, n
