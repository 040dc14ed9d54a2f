"""The learned GJK warm start: a 44-128-128-128-3 MLP whose output seeds GJK.

No counterpart in parry: the seed fills the ``init_dir`` slot of
contact_support_map_support_map_with_params
(src/query/contact/contact_support_map_support_map.rs:52-60), which parry's
own manifold code fills with the previous normal.  The network is trained, by
the library's own bf16 MFMA forward/backward kernels, on the directions the
plain GJK produces (normal1 in shape 1's frame); it only starts the simplex, so
GJK still stops at parry's tolerance, and the plain default seed remains the
fallback (for a degenerate output and when the seeded run hits the iteration cap).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native

NUM_PARAMS = 39171
FEATURE_STRIDE = 48
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_mlp.npy")


def _stream(stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def pair_features(poses, pairs, descriptors, stream=None):
    """(m, 48) f32 CUDA feature rows of shape pairs."""
    import torch

    m = pairs.shape[0]
    out = torch.empty((m, FEATURE_STRIDE), dtype=torch.float32, device=poses.device)
    _native.check(_native.lib().pq_pair_features(poses.data_ptr(), pairs.data_ptr(), m, descriptors.data_ptr(),
                                                 out.data_ptr(), _stream(stream)))
    return out


def gjk_directions(shapes, poses, pairs, stream=None):
    """The plain GJK's normal1 of every pair in shape 1's frame ((m, 3) CUDA f32)
    and a validity mask: query::contact with an unbounded prediction, so every
    pair yields its closest-points (or EPA) direction."""
    import torch

    from .query import contact_pairs

    m = pairs.shape[0]
    c = contact_pairs(shapes, poses, pairs, 3.0e38, stream=stream)
    out = torch.empty((m, 3), dtype=torch.float32, device=poses.device)
    _native.check(_native.lib().pq_contact_local_normals(c.data_ptr(), poses.data_ptr(), pairs.data_ptr(), m,
                                                         out.data_ptr(), _stream(stream)))
    valid = c.view(torch.int32)[:, 13] == 1
    return out, valid


class WarmStartModel:
    """The network on the device (``pq_mlp``)."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        lib = _native.lib()
        h = ctypes.c_void_p()
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            assert p.shape == (NUM_PARAMS,)
            _native.check(lib.pq_mlp_create(p.ctypes.data, seed, ctypes.byref(h)))
        else:
            _native.check(lib.pq_mlp_create(None, seed, ctypes.byref(h)))
        self._h = h

    @classmethod
    def pretrained(cls, path: str = WEIGHTS_FILE) -> "WarmStartModel":
        return cls(np.load(path, allow_pickle=False))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _native.lib().pq_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        p = np.zeros(NUM_PARAMS, np.float32)
        _native.check(_native.lib().pq_mlp_get_params(self._h, p.ctypes.data))
        return p

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        _native.check(_native.lib().pq_mlp_set_params(self._h, p.ctypes.data))

    def forward(self, features, stream=None):
        """Raw outputs (n, 3) for (n, 48) feature rows."""
        import torch

        n = features.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=features.device)
        _native.check(_native.lib().pq_mlp_forward(self._h, features.data_ptr(), n, out.data_ptr(), _stream(stream)))
        return out

    def seeds(self, poses, pairs, descriptors, out=None, stream=None):
        """Unit GJK seeds (m, 3) for shape pairs (zero rows mean parry's default seed)."""
        import torch

        m = pairs.shape[0]
        if out is None:
            out = torch.empty((m, 3), dtype=torch.float32, device=poses.device)
        _native.check(_native.lib().pq_contact_seeds(self._h, poses.data_ptr(), pairs.data_ptr(), m,
                                                     descriptors.data_ptr(), out.data_ptr(), _stream(stream)))
        return out

    def gradients(self, features, labels, rows, stream=None):
        """(loss sum, gradient (NUM_PARAMS,) CUDA f32) over feature rows ``rows``."""
        import torch

        g = torch.empty(NUM_PARAMS, dtype=torch.float32, device=features.device)
        loss = torch.zeros(1, dtype=torch.float64, device=features.device)
        _native.check(_native.lib().pq_mlp_gradients(self._h, features.data_ptr(), labels.data_ptr(), rows.data_ptr(),
                                                     rows.shape[0], g.data_ptr(), loss.data_ptr(), _stream(stream)))
        return float(loss.item()), g

    def train(self, features, labels, epochs: int = 2, batch: int = 8192, lr: float = 2e-3, lr_final: float = 1e-4,
              seed: int = 0, stream=None):
        """Adam on (n, 48) features and (n, 3) unit labels; returns the mean loss per epoch."""
        cfg = _native.MlpTrainConfig(epochs, batch, lr, lr_final, 0.9, 0.999, 1e-8, 0, seed)
        losses = np.zeros(epochs, np.float64)
        _native.check(_native.lib().pq_mlp_train(self._h, features.data_ptr(), labels.data_ptr(), features.shape[0],
                                                 ctypes.byref(cfg), losses.ctypes.data, _stream(stream)))
        return losses


def support_map_pairs(descs, pairs):
    """Mask of the pairs GJK runs on with a seed (no Ball in the pair)."""
    t = np.asarray(descs["type"])
    p = np.asarray(pairs).astype(np.int64)
    return (t[p[:, 0]] != 0) & (t[p[:, 1]] != 0)


def training_set(shapes, descs, poses, pairs, stream=None):
    """(features, labels) CUDA tensors for the seeded (Ball-free) pairs of a scene,
    labelled by the plain GJK."""
    import torch

    keep = torch.from_numpy(support_map_pairs(descs, pairs.cpu().numpy().view(np.uint32))).to(pairs.device)
    sp = pairs[keep].contiguous()
    labels, valid = gjk_directions(shapes, poses, sp, stream)
    sp = sp[valid].contiguous()
    labels = labels[valid].contiguous()
    feats = pair_features(poses, sp, shapes.descriptors(stream=stream), stream)
    return feats, labels
