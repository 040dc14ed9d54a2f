"""A dynamic scene: mixed convex shapes in rigid motion, one collision frame per step.

Each frame runs, on one stream and captured as a hipGraph: shape AABBs at the
current poses (src/shape/shape.rs:369), Bvh refit (insert_or_update_partially +
refit, bvh_insert.rs:209, bvh_refit.rs:170), the overlapping pairs
(traverse_bvtt_single_tree, bvh_traverse_bvtt.rs:19), GJK seeds (the learned
network for every pair, overridden by the previous frame's normal where the
pair persists), query::contact on every pair (contact_shape_shape.rs:123-138),
the cache of this frame's normals for the next, and finally the pose update.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .learned import WarmStartModel
from .partitioning import Bvh
from .query import contact_workspace_bytes
from .shape import Graph, ShapeSet


class DynamicScene:
    """``descs``/``vertices``/``poses`` as from ``scenes.mixed_scene``; velocities
    uniform in [-1, 1]^3 and angular velocities in [-pi, pi]^3 rad/s (seeded).

    ``seeding``: "plain" (parry's default seed everywhere, the default: it is the
    fastest measured on the 262,144-shape scene), "cache" (cached normals of
    persisting pairs, parry's default seed otherwise) or "learned+cache" (network
    seeds, replaced by the cached normal of persisting pairs).  EPA takes most of
    a frame and no GJK seed shortens it, so neither seeded mode pays for its
    extra kernels there.
    """

    def __init__(self, descs, vertices, poses, seed: int = 0, dt: float = 1.0 / 60.0, prediction: float = 0.01,
                 pair_capacity: int | None = None, seeding: str = "plain", model: WarmStartModel | None = None):
        import torch

        self.n = len(descs)
        self.dt = float(dt)
        self.prediction = float(prediction)
        self.seeding = seeding
        self.shapes = ShapeSet(descs, vertices)
        self.stream = torch.cuda.Stream()
        dev = torch.device("cuda")
        rng = np.random.default_rng(seed)
        self.poses = torch.from_numpy(np.ascontiguousarray(poses, np.float32)).to(dev)
        self.linvel = torch.from_numpy(rng.uniform(-1.0, 1.0, (self.n, 3)).astype(np.float32)).to(dev)
        self.angvel = torch.from_numpy(rng.uniform(-np.pi, np.pi, (self.n, 3)).astype(np.float32)).to(dev)
        self.aabbs = torch.empty((self.n, 6), dtype=torch.float32, device=dev)
        self.shapes.compute_aabbs(self.poses, out=self.aabbs)
        torch.cuda.synchronize()
        self.bvh = Bvh(self.aabbs.cpu().numpy())
        self.capacity = int(pair_capacity if pair_capacity is not None else 8 * self.n)
        self.pairs = torch.zeros((self.capacity, 2), dtype=torch.int32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int64, device=dev)
        self.contacts = torch.empty((self.capacity, 14), dtype=torch.float32, device=dev)
        self.seeds = torch.zeros((self.capacity, 3), dtype=torch.float32, device=dev)
        self.workspace = torch.empty(contact_workspace_bytes(self.capacity), dtype=torch.uint8, device=dev)
        lib = _native.lib()
        self.cache_cap = 1 << int(np.ceil(np.log2(4 * self.capacity)))
        cb = int(lib.pq_warm_cache_bytes(self.cache_cap))
        self.caches = [torch.empty(cb, dtype=torch.uint8, device=dev) for _ in range(2)]
        for c in self.caches:
            _native.check(lib.pq_warm_cache_clear(c.data_ptr(), self.cache_cap, None))
        self.model = model if model is not None else (WarmStartModel.pretrained() if "learned" in seeding else None)
        self.desc = self.shapes.descriptors()
        torch.cuda.synchronize()
        self.frame = 0
        self.graphs = [None, None]

    # -- one frame ------------------------------------------------------------
    def _issue(self, parity: int) -> None:
        """The frame's library calls on self.stream (reads cache[parity], writes the other)."""
        lib = _native.lib()
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        chk = _native.check
        chk(lib.pq_shapes_compute_aabbs(self.shapes.handle, self.poses.data_ptr(), self.aabbs.data_ptr(), sp))
        chk(lib.pq_bvh_update_and_refit(self.bvh.handle, self.aabbs.data_ptr(), sp))
        chk(lib.pq_bvh_self_pairs(self.bvh.handle, self.pairs.data_ptr(), self.capacity, self.count.data_ptr(), sp))
        seeds = None
        if self.seeding != "plain":
            seeds = self.seeds.data_ptr()
            if self.model is not None and "learned" in self.seeding:
                chk(lib.pq_contact_seeds_device_count(self.model.handle, self.poses.data_ptr(), self.pairs.data_ptr(),
                                                      self.capacity, self.count.data_ptr(), self.desc.data_ptr(),
                                                      seeds, sp))
            else:  # zero seeds = parry's default seed where no cached normal applies
                chk(lib.pq_device_zero(seeds, self.seeds.numel() * 4, sp))
            chk(lib.pq_warm_cache_apply(self.caches[parity].data_ptr(), self.cache_cap, self.pairs.data_ptr(),
                                        self.count.data_ptr(), self.capacity, seeds, None, sp))
        chk(lib.pq_contact_pairs_device_count(self.shapes.handle, self.poses.data_ptr(), self.pairs.data_ptr(),
                                              self.capacity, self.count.data_ptr(), self.prediction, seeds,
                                              self.contacts.data_ptr(), None, self.workspace.data_ptr(), sp))
        if self.seeding != "plain":
            nxt = self.caches[1 - parity].data_ptr()
            chk(lib.pq_warm_cache_clear(nxt, self.cache_cap, sp))
            chk(lib.pq_warm_cache_store(nxt, self.cache_cap, self.pairs.data_ptr(), self.count.data_ptr(),
                                        self.capacity, self.contacts.data_ptr(), self.poses.data_ptr(), sp))
        chk(lib.pq_integrate_poses(self.poses.data_ptr(), self.linvel.data_ptr(), self.angvel.data_ptr(), self.n,
                                   self.dt, sp))

    def capture(self) -> None:
        """Captures the even and odd frames (they swap the two caches) as hipGraphs."""
        import torch

        lib = _native.lib()
        torch.cuda.synchronize()
        for parity in (0, 1):
            sp = ctypes.c_void_p(self.stream.cuda_stream)
            _native.check(lib.pq_stream_begin_capture(sp))
            try:
                self._issue(parity)
            finally:
                g = ctypes.c_void_p()
                _native.check(lib.pq_stream_end_capture(sp, ctypes.byref(g)))
            self.graphs[parity] = Graph(g, keep=(self,))

    def step(self, use_graph: bool = True) -> None:
        """Runs one frame (asynchronously on self.stream) and advances the poses."""
        parity = self.frame & 1
        if use_graph and self.graphs[parity] is not None:
            self.graphs[parity].launch(self.stream)
        else:
            self._issue(parity)
        self.frame += 1

    def num_pairs(self) -> int:
        """Pairs of the last frame (waits for the scene's stream)."""
        self.stream.synchronize()
        m = int(self.count.item())
        if m > self.capacity:
            raise RuntimeError(f"{m} pairs exceed the capacity {self.capacity}")
        return m
