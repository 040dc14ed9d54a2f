"""Pins the topology of the oracle's binned Bvh build node for node.

The oracle's Bvh::from_leaves(Binned, ..) (bvh_tree.rs:1891-1955 ->
rebuild_range_binned, bvh_binned_build.rs:39-176, then the depth-first refit of
bvh_refit.rs:170-320) decides which triangle a ray reports on ties and which
order pairs come out in, so an off-by-one in its split-plane choice, a swapped
partition step or a wrong axis would change results against parry while every
invariant check still passed.  Here a second implementation, written separately
in plain Python on float32 scalars and organised differently (each range returns
its subtree; nodes are numbered in preorder as they are emitted), rebuilds the
same trees, and the two node arrays must be byte-identical:

  * the reference's own leaf sets: make_test_aabb(i) = the box of half extent 1
    around (i, i, i), for 1..=100 leaves (bvh_tests.rs:5-7, 34-79);
  * random boxes of random sizes, clustered and uniform;
  * degenerate sets: identical boxes (every split falls back to len / 2), two
    clusters of identical centres, centres on a coarse grid (bin ties), a flat
    set (zero extent along two axes);
  * the triangle boxes of the icosphere meshes the ray workloads use.

The product's host builder (parry_amd/csrc/bvh_build.cpp) is checked byte for
byte against the oracle in test_host_abi.py, so this pins it too.
"""
import numpy as np
import pytest

from parry_amd import scenes
from parry_amd.shape import BVH_NODE_WIDE_DTYPE

F = np.float32
NUM_BINS = 8
BIN_EPSILON = F(1.0e-5)
CHANGED = 1 << 30          # bvh_tree.rs:157, after refit resolves a pending change
CHANGE_PENDING = 3 << 30   # bvh_tree.rs:158, the state BvhNode::leaf starts in


def _as_usize(x):
    """Rust `f as usize`: NaN and negatives give 0, large values saturate."""
    if not (x > 0):
        return 0
    if not np.isfinite(x) or x >= F(2.0 ** 64):
        return 2 ** 64 - 1
    return int(x)


class _Leaf:
    __slots__ = ("lo", "hi", "id")

    def __init__(self, lo, hi, i):
        self.lo, self.hi, self.id = lo, hi, i

    def centre(self, k):
        return (self.lo[k] + self.hi[k]) * F(0.5)


def _volume(lo, hi):
    return ((hi[0] - lo[0]) * (hi[1] - lo[1])) * (hi[2] - lo[2])


def _split(leaves):
    """Orders `leaves` in place as parry's partition does and returns the split index."""
    cs = [[lf.centre(k) for k in range(3)] for lf in leaves]
    lo = [min(c[k] for c in cs) for k in range(3)]
    hi = [max(c[k] for c in cs) for k in range(3)]
    ext = [hi[k] - lo[k] for k in range(3)]
    axis = 0  # nalgebra's imax: the first of equal maxima
    for k in (1, 2):
        if ext[k] > ext[axis]:
            axis = k
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        k1 = F(NUM_BINS) * (F(1.0) - BIN_EPSILON) / (hi[axis] - lo[axis])
        k0 = lo[axis]
        bin_of = [_as_usize(k1 * (lf.centre(axis) - k0)) for lf in leaves]
    assert max(bin_of) < NUM_BINS

    big = F(3.4028235e38)
    counts = [0] * NUM_BINS
    boxes = [([big] * 3, [-big] * 3) for _ in range(NUM_BINS)]
    for lf, b in zip(leaves, bin_of):
        counts[b] += 1
        blo, bhi = boxes[b]
        boxes[b] = ([min(blo[k], lf.lo[k]) for k in range(3)], [max(bhi[k], lf.hi[k]) for k in range(3)])

    # suffix merges of bins i..7 for i >= 1, and the prefix merge grown plane by plane
    suffix = {}
    acc_lo, acc_hi, acc_n = boxes[-1][0], boxes[-1][1], counts[-1]
    suffix[NUM_BINS - 1] = (acc_lo, acc_hi, acc_n)
    for b in range(NUM_BINS - 2, 0, -1):
        acc_lo = [min(acc_lo[k], boxes[b][0][k]) for k in range(3)]
        acc_hi = [max(acc_hi[k], boxes[b][1][k]) for k in range(3)]
        acc_n += counts[b]
        suffix[b] = (acc_lo, acc_hi, acc_n)

    best_cost, best_plane, best_left = big, 0, counts[0]
    pre_lo, pre_hi, pre_n = boxes[0][0], boxes[0][1], counts[0]
    with np.errstate(over="ignore", invalid="ignore"):
        for plane in range(NUM_BINS - 1):
            r_lo, r_hi, r_n = suffix[plane + 1]
            cost = _volume(pre_lo, pre_hi) * F(pre_n) + _volume(r_lo, r_hi) * F(r_n)
            if cost < best_cost:
                best_cost, best_plane, best_left = cost, plane, pre_n
            pre_lo = [min(pre_lo[k], boxes[plane + 1][0][k]) for k in range(3)]
            pre_hi = [max(pre_hi[k], boxes[plane + 1][1][k]) for k in range(3)]
            pre_n += counts[plane + 1]

    n = len(leaves)
    mid = best_left
    if mid == 0 or mid == n:
        return n // 2
    # two cursors: a left leaf beyond the plane swaps with a right leaf before it
    i, j = 0, mid
    while i != mid and j != n:
        while bin_of[i] <= best_plane:
            i += 1
            if i == mid:
                return mid
        while bin_of[j] > best_plane:
            j += 1
            if j == n:
                return mid
        leaves[i], leaves[j] = leaves[j], leaves[i]
        bin_of[i], bin_of[j] = bin_of[j], bin_of[i]
        i += 1
        j += 1
    return mid


def _emit(leaves, rows):
    """Emits the subtree over `leaves` (>= 2) in preorder; returns (row, lo, hi, count)."""
    me = len(rows)
    rows.append(None)
    mid = _split(leaves)
    sides = []
    for part in (leaves[:mid], leaves[mid:]):
        if len(part) == 1:
            lf = part[0]
            sides.append((lf.id, lf.lo, lf.hi, 1))
        else:
            sides.append(_emit(part, rows))
    (lc, llo, lhi, ln), (rc, rlo, rhi, rn) = sides
    rows[me] = (lc, llo, lhi, ln, rc, rlo, rhi, rn)
    lo = [min(llo[k], rlo[k]) for k in range(3)]
    hi = [max(lhi[k], rhi[k]) for k in range(3)]
    return me, lo, hi, ln + rn


def binned_nodes(aabbs) -> np.ndarray:
    a = np.asarray(aabbs, np.float32).reshape(-1, 6)
    leaves = [_Leaf([F(x) for x in a[i, :3]], [F(x) for x in a[i, 3:]], i) for i in range(a.shape[0])]
    out = np.zeros(max(1, len(leaves) - 1), BVH_NODE_WIDE_DTYPE)
    if len(leaves) <= 2:  # bvh_tree.rs:1914-1930: no rebuild and no refit, the leaves stay pending
        for side, lf in zip(("left", "right"), leaves):
            out[0][side + "_mins"], out[0][side + "_maxs"] = lf.lo, lf.hi
            out[0][side + "_children"], out[0][side + "_data"] = lf.id, 1 | CHANGE_PENDING
        return out
    rows = []
    _emit(leaves, rows)
    assert len(rows) == out.shape[0]
    for r, (lc, llo, lhi, ln, rc, rlo, rhi, rn) in enumerate(rows):
        out[r]["left_mins"], out[r]["left_maxs"], out[r]["left_children"] = llo, lhi, lc
        out[r]["right_mins"], out[r]["right_maxs"], out[r]["right_children"] = rlo, rhi, rc
        out[r]["left_data"], out[r]["right_data"] = ln | CHANGED, rn | CHANGED
    return out


def _same(orc, aabbs, label):
    got = orc.OracleBvh(aabbs).nodes()
    want = binned_nodes(aabbs)
    assert got.shape == want.shape, label
    diff = np.nonzero(got.view(np.uint8).reshape(got.shape[0], -1) != want.view(np.uint8).reshape(want.shape[0], -1))
    assert diff[0].size == 0, f"{label}: first differing node {diff[0][0]}"


def _boxes(centres, half):
    c = np.asarray(centres, np.float32)
    h = np.broadcast_to(np.asarray(half, np.float32), c.shape)
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def test_reference_diagonal_leaf_sets(orc):
    # bvh_tests.rs:5-7: Aabb::from_half_extents((i, i, i), (1, 1, 1))
    for n in range(1, 101):
        i = np.arange(n, dtype=np.float32)[:, None].repeat(3, 1)
        _same(orc, _boxes(i, 1.0), f"diagonal {n}")


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (7, 3), (33, 4), (100, 5), (1000, 6), (3000, 7)])
def test_random_boxes(orc, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (n, 3)) * rng.uniform(0.1, 1.0, 3)
    _same(orc, _boxes(c, rng.uniform(0.01, 2.0, (n, 3))), f"uniform {n}")
    cl = rng.normal(0, 1, (n, 3)) + rng.integers(0, 4, (n, 1)) * 7.0
    _same(orc, _boxes(cl, rng.uniform(0.0, 0.5, (n, 3))), f"clusters {n}")


def test_degenerate_sets(orc):
    rng = np.random.default_rng(9)
    _same(orc, _boxes(np.zeros((50, 3)), 1.0), "identical boxes")
    two = np.repeat(np.array([[0, 0, 0], [3, 1, 0]], np.float32), [37, 11], axis=0)
    _same(orc, _boxes(two, rng.uniform(0.1, 1.0, (48, 3))), "two centres")
    grid = rng.integers(0, 3, (400, 3)).astype(np.float32)
    _same(orc, _boxes(grid, 0.25), "coarse grid")
    flat = np.stack([rng.uniform(0, 5, 300), np.zeros(300), np.zeros(300)], 1)
    _same(orc, _boxes(flat, rng.uniform(0.1, 1.0, (300, 3))), "flat")


@pytest.mark.parametrize("level", [1, 3, 4])
def test_icosphere_triangle_boxes(orc, level):
    v, f = scenes.noisy_icosphere(level, seed=1)
    t = v[f]
    aabbs = np.concatenate([t.min(1), t.max(1)], 1).astype(np.float32)
    _same(orc, aabbs, f"icosphere {level}")
    # the TriMesh's own tree is this tree
    assert np.array_equal(orc.OracleTriMesh(v, f).bvh_nodes().view(np.uint8), binned_nodes(aabbs).view(np.uint8))
