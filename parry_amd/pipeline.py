"""One collision frame on the GPU: leaf AABBs -> Bvh refit -> pairs -> contacts.

This strings together the parry3d calls a physics step makes for a set of
posed shapes (src/shape/shape.rs:369 compute_aabb, bvh_insert.rs:209
insert_or_update_partially, bvh_refit.rs:170 refit,
bvh_traverse_bvtt.rs:19 traverse_bvtt_single_tree, then
contact_shape_shape.rs:123 query::contact on every reported pair), with every
buffer preallocated and the pair count kept on the device, so ``step`` issues
only asynchronous launches on one stream (capturable in a hipGraph).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .partitioning import Bvh
from .query import contact_workspace_bytes


class CollisionPipeline:
    def __init__(self, shapes, poses, prediction: float = 0.01, pair_capacity: int | None = None, stream=None):
        import torch

        self.shapes = shapes
        self.prediction = float(prediction)
        self.device = poses.device
        n = len(shapes)
        self.aabbs = torch.empty((n, 6), dtype=torch.float32, device=self.device)
        shapes.compute_aabbs(poses, out=self.aabbs, stream=stream)
        torch.cuda.synchronize(self.device)
        self.bvh = Bvh(self.aabbs.cpu().numpy())  # Bvh::from_leaves(Binned, ..) on the host, once
        self.capacity = int(pair_capacity if pair_capacity is not None else max(64, 16 * n))
        self.pairs = torch.empty((self.capacity, 2), dtype=torch.int32, device=self.device)
        self.count = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.contacts = torch.empty((self.capacity, 14), dtype=torch.float32, device=self.device)
        self.iters = torch.empty((self.capacity, 2), dtype=torch.int32, device=self.device)
        self.workspace = torch.empty(contact_workspace_bytes(self.capacity), dtype=torch.uint8, device=self.device)

    def _stream(self, stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def step(self, poses, init_dirs=None, with_iters: bool = False, stream=None) -> None:
        """One frame; results stay on the device (``num_pairs()``, ``pairs``, ``contacts``)."""
        lib = _native.lib()
        sp = self._stream(stream)
        _native.check(lib.pq_shapes_compute_aabbs(self.shapes.handle, poses.data_ptr(), self.aabbs.data_ptr(), sp))
        _native.check(lib.pq_bvh_update_and_refit(self.bvh.handle, self.aabbs.data_ptr(), sp))
        _native.check(lib.pq_bvh_self_pairs(self.bvh.handle, self.pairs.data_ptr(), self.capacity,
                                            self.count.data_ptr(), sp))
        _native.check(lib.pq_contact_pairs_device_count(
            self.shapes.handle, poses.data_ptr(), self.pairs.data_ptr(), self.capacity, self.count.data_ptr(),
            self.prediction, None if init_dirs is None else init_dirs.data_ptr(), self.contacts.data_ptr(),
            self.iters.data_ptr() if with_iters else None, self.workspace.data_ptr(), sp))

    def num_pairs(self) -> int:
        """Pairs reported by the last step (synchronises); raises if they overflowed the capacity."""
        m = int(self.count.item())
        if m > self.capacity:
            raise RuntimeError(f"{m} pairs exceed the pipeline's capacity {self.capacity}")
        return m

    def results(self):
        """(pairs (m, 2) int32, contacts (m, 14) f32) of the last step, on the device."""
        m = self.num_pairs()
        return self.pairs[:m], self.contacts[:m]
