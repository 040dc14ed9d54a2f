"""pq_contact_pairs_noalloc (the contact entry point that allocates nothing)
against pq_contact_pairs, through the C ABI.

Bar: records and GJK/EPA iteration counts byte-identical to pq_contact_pairs on
the same inputs (both run the same kernels), on two calls in a row, the second
with fewer pairs than the first, so that anything the first call left in the
fixed workspace would show.  Sizes above PQ_CONTACT_NOALLOC_MAX_PAIRS are
rejected on the host.
"""
import ctypes

import numpy as np
import pytest

from parry_amd import Bvh, ShapeSet, scenes

pytestmark = pytest.mark.gpu
MAX_PAIRS = 262144


def _call(entry, shapes, poses, pairs, prediction, workspace=False):
    import torch

    from parry_amd import _native

    m = pairs.shape[0]
    out = torch.full((m, 14), float("nan"), dtype=torch.float32, device="cuda")
    iters = torch.full((m, 2), -1, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _native.lib()
    if entry == "noalloc":
        rc = lib.pq_contact_pairs_noalloc(shapes.handle, poses.data_ptr(), pairs.data_ptr(), m, prediction, None,
                                          out.data_ptr(), iters.data_ptr(), stream)
    else:
        rc = lib.pq_contact_pairs(shapes.handle, poses.data_ptr(), pairs.data_ptr(), m, prediction, None,
                                  out.data_ptr(), iters.data_ptr(), None, stream)
    _native.check(rc)
    torch.cuda.synchronize()
    return out.cpu().numpy(), iters.cpu().numpy()


def _scene_pairs(n, seed, limit):
    import torch

    sc = scenes.mixed_scene(n, seed=seed)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    pairs = Bvh(shapes.compute_aabbs(poses).cpu().numpy()).self_pairs()
    return shapes, poses, pairs[:limit].contiguous()


def test_noalloc_matches_contact_pairs_on_two_calls():
    # first call: up to the maximum; second call: a different, smaller scene
    cases = [(60000, 11, MAX_PAIRS), (3000, 12, 5000)]
    for n, seed, limit in cases:
        shapes, poses, pairs = _scene_pairs(n, seed, limit)
        m = pairs.shape[0]
        print(f"case: noalloc vs pq_contact_pairs, mixed_scene({n}, {seed}), {m} pairs", flush=True)
        assert 0 < m <= MAX_PAIRS
        got, got_it = _call("noalloc", shapes, poses, pairs, 0.01)
        ref, ref_it = _call("workspace", shapes, poses, pairs, 0.01)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), "records differ"
        assert np.array_equal(got_it, ref_it), "iteration counts differ"
        print(f"{m} pairs: {int((got.view(np.uint32)[:, 13] == 1).sum())} contacts, "
              f"{int((got_it[:, 1] > 0).sum())} through EPA, byte-identical", flush=True)


def test_noalloc_rejects_too_many_pairs():
    import torch

    from parry_amd import _native

    sc = scenes.mixed_scene(64, seed=1)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    lib = _native.lib()
    rc = lib.pq_contact_pairs_noalloc(shapes.handle, None, None, MAX_PAIRS + 1, 0.01, None, None, None, None)
    assert rc == _native.PQ_ERR_INVALID
    assert "maximum" in _native.last_error()
    assert lib.pq_contact_pairs_noalloc(shapes.handle, None, None, 0, 0.01, None, None, None, None) == _native.PQ_OK
    torch.cuda.synchronize()
