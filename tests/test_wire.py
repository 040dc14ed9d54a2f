"""The serialized Bvh / TriMesh (host only, no GPU).

parry derives serde for Bvh {nodes, parents, leaf_node_indices} (bvh_tree.rs:1748-1762)
and TriMesh {bvh, vertices, indices, pseudo_normals, topology, connected_components,
flags} (trimesh.rs:510-528); the library writes them as bincode 1.x's default options
lay them out.  Pinned by:

  * an encoder written here from that field list with ``struct`` (independent of the
    C++ one), byte for byte, on the reference's diagonal leaf sets and random trees,
    and a known-answer byte string for a three-leaf tree assembled by hand;
  * parents and leaf_node_indices recomputed from the nodes by walking the tree
    (parents[w] = the slot holding wide node w, leaf_node_indices[i] = the slot
    holding leaf i, slots as BvhNodeIndex::new(wide, is_right), bvh_tree.rs:1403-1414);
  * round trips, and rejection of truncated, inconsistent and unsupported inputs.

Parity against bytes written by parry itself is unpinned: parry3d (Rust, serde,
bincode) cannot be built here.
"""
import ctypes
import struct

import numpy as np
import pytest

from parry_amd import _native, scenes
from parry_amd.shape import BVH_NODE_WIDE_DTYPE
from parry_amd.wire import bvh_from_bytes, bvh_to_bytes, trimesh_from_bytes, trimesh_to_bytes

LC = 0x3FFFFFFF


def _binned_nodes(aabbs):
    lib = _native.lib()
    a = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
    n, d = ctypes.c_uint32(), ctypes.c_uint32()
    lib.pq_bvh_from_leaves(a.ctypes.data, a.shape[0], None, 0, ctypes.byref(n), ctypes.byref(d))
    out = np.zeros(n.value, BVH_NODE_WIDE_DTYPE)
    _native.check(lib.pq_bvh_from_leaves(a.ctypes.data, a.shape[0], out.ctypes.data, out.shape[0], ctypes.byref(n),
                                         ctypes.byref(d)))
    return out


def _diagonal(n):
    c = np.arange(n, dtype=np.float32)[:, None].repeat(3, 1)
    return np.concatenate([c - 1, c + 1], 1)


def _slots_by_walk(nodes):
    """parents and leaf slots found by walking down from the root."""
    parents = {0: 0}
    leaf = {}
    stack = [0]
    while stack:
        w = stack.pop()
        for s, side in enumerate(("left", "right")):
            lc = int(nodes[w][side + "_data"]) & LC
            c = int(nodes[w][side + "_children"])
            if lc == 1:
                leaf[c] = (w << 1) | s
            elif lc > 1:
                parents[c] = (w << 1) | s
                stack.append(c)
    return [parents[w] for w in range(nodes.shape[0])], [leaf[i] for i in range(len(leaf))]


def _bincode_bvh(nodes):
    parents, leaf = _slots_by_walk(nodes) if nodes.shape[0] else ([], [])
    out = struct.pack("<Q", nodes.shape[0]) + nodes.tobytes()
    out += struct.pack("<Q", len(parents)) + b"".join(struct.pack("<Q", p) for p in parents)
    out += struct.pack("<QQ", len(leaf), len(leaf)) + b"".join(struct.pack("<BQ", 1, s) for s in leaf)
    return out


def test_three_leaf_known_answer():
    # Three diagonal leaves: the root holds one leaf and a wide node over the
    # other two; the expected bytes are assembled field by field from the nodes.
    nodes = _binned_nodes(_diagonal(3))
    assert nodes.shape[0] == 2
    got = bvh_to_bytes(nodes)
    r = nodes[0]
    inner = 0 if (int(r["left_data"]) & LC) > 1 else 1  # BvhNodeIndex::left(0) or right(0)
    expect = bytearray()
    expect += (2).to_bytes(8, "little")
    expect += nodes.tobytes()
    expect += (2).to_bytes(8, "little") + (0).to_bytes(8, "little") + (inner).to_bytes(8, "little")
    expect += (3).to_bytes(8, "little") + (3).to_bytes(8, "little")
    slots = {}
    for w in range(2):
        for s, side in enumerate(("left", "right")):
            if (int(nodes[w][side + "_data"]) & LC) == 1:
                slots[int(nodes[w][side + "_children"])] = 2 * w + s
    for i in range(3):
        expect += b"\x01" + slots[i].to_bytes(8, "little")
    assert len(got) == 8 + 128 + 8 + 16 + 16 + 27
    assert got == bytes(expect)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 31, 100])
def test_reference_diagonal_sets_match_independent_encoder(n):
    nodes = _binned_nodes(_diagonal(n))
    got = bvh_to_bytes(nodes)
    assert got == _bincode_bvh(nodes)
    assert bvh_from_bytes(got).tobytes() == nodes.tobytes()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_trees_round_trip(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 3000))
    c = rng.uniform(-10, 10, (n, 3))
    aabbs = np.concatenate([c - rng.uniform(0, 1, (n, 3)), c + rng.uniform(0, 1, (n, 3))], 1)
    nodes = _binned_nodes(aabbs)
    data = bvh_to_bytes(nodes)
    assert data == _bincode_bvh(nodes)
    assert len(data) == 8 + 64 * nodes.shape[0] + 8 + 8 * nodes.shape[0] + 16 + 9 * n
    assert bvh_from_bytes(data).tobytes() == nodes.tobytes()


def test_trimesh_layout_and_round_trip():
    v, f = scenes.noisy_icosphere(3, seed=5)
    nodes = _binned_nodes(np.concatenate([v[f].min(1), v[f].max(1)], 1))
    data = trimesh_to_bytes(v, f, nodes)
    bvh = _bincode_bvh(nodes)
    expect = bvh + struct.pack("<Q", v.shape[0]) + v.astype("<f4").tobytes()
    expect += struct.pack("<Q", f.shape[0]) + f.astype("<u4").tobytes() + b"\x00\x00\x00" + struct.pack("<H", 0)
    assert data == expect
    v2, f2, n2 = trimesh_from_bytes(data)
    assert np.array_equal(v2, v) and np.array_equal(f2, f) and n2.tobytes() == nodes.tobytes()


def test_malformed_inputs_are_rejected():
    nodes = _binned_nodes(_diagonal(20))
    data = bvh_to_bytes(nodes)
    with pytest.raises(_native.NativeError):
        bvh_from_bytes(data[:-3])  # truncated
    with pytest.raises(_native.NativeError):
        bvh_from_bytes(data + b"\x00")  # trailing bytes
    bad = bytearray(data)
    off = 8 + 64 * nodes.shape[0] + 8 + 8  # parents[1]
    bad[off] ^= 2
    with pytest.raises(_native.NativeError):
        bvh_from_bytes(bytes(bad))  # parents disagree with the nodes
    v, f = scenes.noisy_icosphere(1, seed=1)
    tn = _binned_nodes(np.concatenate([v[f].min(1), v[f].max(1)], 1))
    mesh = bytearray(trimesh_to_bytes(v, f, tn))
    mesh[-3] = 1  # topology: Some(..)
    with pytest.raises(_native.NativeError) as e:
        trimesh_from_bytes(bytes(mesh))
    assert e.value.code == _native.PQ_ERR_UNSUPPORTED
    # a tree whose leaves are not the triangles
    with pytest.raises(_native.NativeError):
        trimesh_from_bytes(trimesh_to_bytes(v, f[:-1], tn))
