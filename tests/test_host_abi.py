"""CPU-side checks of the product library: the C ABI exports and the host Bvh builder.

The builder (parry_amd/csrc/bvh_build.cpp) must produce parry's binned tree
byte for byte, because the kernels' tie order follows the tree; it is compared
with the oracle's restatement of bvh_binned_build.rs + bvh_refit.rs on meshes,
on the reference's test mesh, and on degenerate leaf sets that exercise the
len/2 fallback split and saturated bin indices.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from parry_amd import BVH_NODE_WIDE_DTYPE, _native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|void|uint32_t)\s*\*?\s*(pq_\w+)\s*\(", text, flags=re.M):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol(native):
    syms = declared_symbols()
    assert len(syms) >= 10
    missing = [s for s in sorted(syms) if not hasattr(native, s)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert syms <= set(_native.SIGNATURES), sorted(syms - set(_native.SIGNATURES))


def build_product(aabbs: np.ndarray):
    a = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
    n = ctypes.c_uint32()
    d = ctypes.c_uint32()
    lib = _native.lib()
    rc = lib.pq_bvh_from_leaves(a.ctypes.data, a.shape[0], None, 0, ctypes.byref(n), ctypes.byref(d))
    if a.shape[0] == 0:
        assert rc == 0 and n.value == 0
        return np.zeros(0, BVH_NODE_WIDE_DTYPE), 0
    assert rc == _native.PQ_ERR_CAPACITY
    out = np.zeros(n.value, BVH_NODE_WIDE_DTYPE)
    _native.check(lib.pq_bvh_from_leaves(a.ctypes.data, a.shape[0], out.ctypes.data, out.shape[0], ctypes.byref(n), ctypes.byref(d)))
    return out, d.value


def tri_aabbs(v, f):
    t = v[f]
    return np.concatenate([t.min(1), t.max(1)], 1).astype(np.float32)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4, 5, 6])
def test_builder_matches_oracle_icosphere(orc, level):
    v, f = scenes.noisy_icosphere(level, seed=level + 10)
    ref = orc.OracleTriMesh(v, f)
    out, depth = build_product(tri_aabbs(v, f))
    assert out.tobytes() == ref.bvh_nodes().tobytes()
    assert depth == ref.depth()


def test_builder_matches_oracle_bunny(orc):
    v, f = scenes.load_obj(os.path.join(ROOT, "tests", "golden", "low_poly_bunny.obj"))
    ref = orc.OracleTriMesh(v, f)
    out, _ = build_product(tri_aabbs(v, f))
    assert out.tobytes() == ref.bvh_nodes().tobytes()


def _oracle_tree_from_aabbs(orc, aabbs):
    # A degenerate triangle (a, b, b) has exactly the AABB spanned by a and b.
    n = aabbs.shape[0]
    v = np.zeros((2 * n, 3), np.float32)
    v[0::2] = aabbs[:, :3]
    v[1::2] = aabbs[:, 3:]
    f = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    return orc.OracleTriMesh(v, f).bvh_nodes()


@pytest.mark.parametrize(
    "case",
    ["identical", "collinear", "two_clusters", "tiny", "one", "two", "three", "random"],
)
def test_builder_matches_oracle_degenerate(orc, case):
    rng = np.random.default_rng(5)
    if case == "identical":  # every centre in one bin -> len/2 fallback at every level
        a = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (37, 1))
    elif case == "collinear":
        x = np.arange(50, dtype=np.float32)
        a = np.stack([x, 0 * x, 0 * x, x + 1, 0 * x + 1, 0 * x + 1], 1)
    elif case == "two_clusters":
        c = np.concatenate([np.zeros(20), np.full(13, 1e6)]).astype(np.float32)
        a = np.stack([c, c, c, c + 1, c + 1, c + 1], 1)
    elif case == "tiny":
        c = (np.arange(40, dtype=np.float32) * np.float32(1e-30))
        a = np.stack([c, c, c, c, c, c], 1)
    elif case == "one":
        a = np.array([[0, 0, 0, 1, 1, 1]], np.float32)
    elif case == "two":
        a = np.array([[0, 0, 0, 1, 1, 1], [2, 0, 0, 3, 1, 1]], np.float32)
    elif case == "three":
        a = np.array([[0, 0, 0, 1, 1, 1], [2, 0, 0, 3, 1, 1], [4, 0, 0, 5, 1, 1]], np.float32)
    else:
        lo = rng.uniform(-10, 10, (500, 3)).astype(np.float32)
        a = np.concatenate([lo, lo + rng.uniform(0, 2, (500, 3)).astype(np.float32)], 1)
    out, _ = build_product(a)
    ref = _oracle_tree_from_aabbs(orc, a.astype(np.float32))
    assert out.tobytes() == ref.tobytes()


def test_builder_empty():
    out, depth = build_product(np.zeros((0, 6), np.float32))
    assert out.shape[0] == 0 and depth == 0


def test_error_reporting(native):
    # Invalid arguments come back as status codes with a message, never a crash.
    h = ctypes.c_void_p()
    v = np.zeros((3, 3), np.float32)
    i = np.array([[0, 1, 5]], np.uint32)  # out of range vertex index
    rc = native.pq_trimesh_create(v.ctypes.data, 3, i.ctypes.data, 1, None, 0, ctypes.byref(h))
    assert rc == _native.PQ_ERR_INVALID
    assert "out of range" in _native.last_error()
    rc = native.pq_trimesh_create(v.ctypes.data, 3, i.ctypes.data, 0, None, 0, ctypes.byref(h))
    assert rc == _native.PQ_ERR_INVALID and "empty index" in _native.last_error()


def test_polyhedron_params_rejected(native):
    # A ConvexPolyhedron record keeps its face range in params[0..1]; only
    # pq_shapes_create's zero params plus pq_shapes_set_polyhedron_faces may set
    # it, so nonzero params are refused before anything reaches the device.
    from parry_amd import SHAPE_DESC_DTYPE

    d = np.zeros(1, SHAPE_DESC_DTYPE)
    d["type"], d["first_vertex"], d["num_vertices"] = 4, 0, 4
    d["params"] = (0, 3, 0, 0)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    h = ctypes.c_void_p()
    rc = native.pq_shapes_create(d.ctypes.data, 1, v.ctypes.data, 4, ctypes.byref(h))
    assert rc == _native.PQ_ERR_INVALID and "zero params" in _native.last_error()
    assert not h.value
