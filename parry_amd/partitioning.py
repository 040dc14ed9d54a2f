"""Broadphase Bvh backed by device memory (parry3d ``src/partitioning/bvh``).

``Bvh`` mirrors the parry3d methods on the broadphase path:
``Bvh::from_leaves`` (bvh_tree.rs:1835), ``insert_or_update_partially`` for every
leaf followed by ``refit`` (bvh_insert.rs:209, bvh_refit.rs:170), and
``traverse_bvtt_single_tree::<false>`` (bvh_traverse_bvtt.rs:19) returning the
overlapping leaf pairs as an array instead of a callback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .shape import BVH_NODE_WIDE_DTYPE


def _stream_ptr(stream, device):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream(device)
    return ctypes.c_void_p(s.cuda_stream)


class Bvh:
    """A device-resident Bvh over ``n`` leaf AABBs (n x 6 float32: mins, maxs)."""

    def __init__(self, aabbs):
        a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_bvh_create(a.ctypes.data, a.shape[0], ctypes.byref(h)))
        self._h = h
        nl, nn, d = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(_native.lib().pq_bvh_info(h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(d)))
        self.num_leaves, self.num_nodes, self.depth = nl.value, nn.value, d.value
        self._count = None

    @classmethod
    def from_leaves(cls, aabbs, strategy: str = "binned", stream=None) -> "Bvh":
        """``Bvh::from_leaves(strategy, aabbs)`` (bvh_tree.rs:1835): "binned" builds on
        the host from an array; "ploc" (BvhBuildStrategy::Ploc, bvh_ploc_build.rs:10-93)
        builds on the GPU from a CUDA tensor (n, 6) f32."""
        if strategy == "binned":
            return cls(aabbs)
        if strategy != "ploc":
            raise ValueError(f"unknown build strategy {strategy!r}")
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_bvh_create_ploc(aabbs.data_ptr(), aabbs.shape[0], ctypes.byref(h),
                                                       _stream_ptr(stream, aabbs.device)))
        self._h = h
        nl, nn, d = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(_native.lib().pq_bvh_info(h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(d)))
        self.num_leaves, self.num_nodes, self.depth = nl.value, nn.value, d.value
        self._count = None
        return self

    @classmethod
    def from_nodes(cls, nodes, num_leaves: int) -> "Bvh":
        """A broadphase handle over existing depth-first wide records (pq_bvh_create_from_nodes)."""
        n = np.ascontiguousarray(nodes, dtype=BVH_NODE_WIDE_DTYPE).reshape(-1)
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        _native.check(_native.lib().pq_bvh_create_from_nodes(n.ctypes.data if n.size else None, n.shape[0],
                                                             int(num_leaves), ctypes.byref(h)))
        self._h = h
        nl, nn, d = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _native.check(_native.lib().pq_bvh_info(h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(d)))
        self.num_leaves, self.num_nodes, self.depth = nl.value, nn.value, d.value
        self._count = None
        return self

    def to_bytes(self) -> bytes:
        """parry's serialized ``Bvh`` (serde as bincode 1.x writes it, bvh_tree.rs:1748-1762) of the current tree."""
        from .wire import bvh_to_bytes

        return bvh_to_bytes(self.nodes())

    @classmethod
    def from_bytes(cls, data: bytes) -> "Bvh":
        from .wire import bvh_from_bytes

        nodes = bvh_from_bytes(data)
        leaves = sum(int(((nodes[s + "_data"] & 0x3FFFFFFF) == 1).sum()) for s in ("left", "right"))
        return cls.from_nodes(nodes, leaves)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_bvh_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def nodes(self) -> np.ndarray:
        """The current tree as depth-first BvhNodeWide records (synchronises)."""
        out = np.zeros(self.num_nodes, dtype=BVH_NODE_WIDE_DTYPE)
        if self.num_nodes:
            _native.check(_native.lib().pq_bvh_export(self._h, out.ctypes.data, out.shape[0]))
        return out

    def update_and_refit(self, aabbs, stream=None) -> None:
        """insert_or_update_partially(aabbs[i], i, 0.0) for all i, then refit (CUDA tensor (n, 6) f32)."""
        _native.check(_native.lib().pq_bvh_update_and_refit(self._h, aabbs.data_ptr(), _stream_ptr(stream, aabbs.device)))

    def self_pairs(self, capacity: int | None = None, out=None, stream=None, device="cuda"):
        """All overlapping leaf pairs (i < j) as a CUDA int32 tensor (m, 2).

        Runs the counting and writing passes; if ``capacity`` is too small the
        call is repeated with the exact size (one host synchronisation).
        """
        import torch

        dev = torch.device(device)
        if self._count is None:
            self._count = torch.zeros(1, dtype=torch.int64, device=dev)
        cap = capacity if capacity is not None else max(16, 8 * self.num_leaves)
        if out is None or out.numel() < 2 * cap:
            out = torch.empty((cap, 2), dtype=torch.int32, device=dev)
        sp = _stream_ptr(stream, dev)
        lib = _native.lib()
        _native.check(lib.pq_bvh_self_pairs(self._h, out.data_ptr(), cap, self._count.data_ptr(), sp))
        m = int(self._count.item())
        if m > cap:
            out = torch.empty((m, 2), dtype=torch.int32, device=dev)
            _native.check(lib.pq_bvh_self_pairs(self._h, out.data_ptr(), m, self._count.data_ptr(), sp))
        return out[:m]


def ploc_nodes(aabbs, out=None, workspace=None, stream=None):
    """``Bvh::from_leaves(BvhBuildStrategy::Ploc, aabbs)``'s depth-first wide nodes built
    on the GPU (pq_bvh_from_leaves_ploc): ``aabbs`` a CUDA tensor (n, 6) f32.  Returns
    (nodes uint8 CUDA tensor of n - 1 (1 for n <= 2) 64-byte records, depth); view
    them with ``BVH_NODE_WIDE_DTYPE`` after a copy to the host.  Synchronises ``stream``."""
    import torch

    n = aabbs.shape[0]
    if not (aabbs.is_cuda and aabbs.dtype == torch.float32 and aabbs.is_contiguous() and aabbs.shape[-1] == 6):
        raise ValueError("ploc_nodes: aabbs must be a contiguous CUDA float32 tensor (n, 6)")
    nn = 0 if n == 0 else max(1, n - 1)
    if out is None:
        out = torch.empty((max(nn, 1), 64), dtype=torch.uint8, device=aabbs.device)
    got, depth = ctypes.c_uint32(), ctypes.c_uint32()
    _native.check(_native.lib().pq_bvh_from_leaves_ploc(
        aabbs.data_ptr(), n, out.data_ptr(), out.shape[0], ctypes.byref(got), ctypes.byref(depth),
        None if workspace is None else workspace.data_ptr(), _stream_ptr(stream, aabbs.device)))
    return out[:got.value], depth.value
