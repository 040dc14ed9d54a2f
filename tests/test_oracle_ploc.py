"""Oracle pins for the PLOC Bvh build (Bvh::from_leaves(BvhBuildStrategy::Ploc),
bvh_tree.rs:1891-1955 with rebuild_range_ploc, bvh_ploc_build.rs:10-93, and the
Morton code of utils/morton.rs:13-40).

  * the reference's bvh_build_and_removal test: the PLOC trees of its diagonal
    leaves (centres (i, i, i), half extents 1) for 1..=100 leaves are well formed
    (Bvh::assert_well_formed);
  * Morton codes against a bit-by-bit interleave (bit b of x, y, z at 3b, 3b+1,
    3b+2), including the 2^21 wrap of a coordinate of exactly 1.0 and the
    saturating casts of NaN and negative inputs;
  * the whole tree against a second, independent implementation of the
    algorithm in plain Python/numpy float32 (key sort, neighbour search, mutual
    merges, depth-first relayout), node for node, on random leaf sets;
  * the tree serves queries: its self-pairs equal the brute-force overlap set.
"""
import numpy as np
import pytest

from bvh_checks import check_tree
from oracle import oracle as orc

LC = 0x3FFFFFFF


def _diag(n):
    c = np.arange(n, dtype=np.float32)[:, None].repeat(3, 1)
    return np.concatenate([c - 1, c + 1], 1).astype(np.float32)


@pytest.mark.parametrize("n", list(range(1, 101)))
def test_reference_diagonal_leaves_well_formed(n):
    nodes = orc.OracleBvh(_diag(n), ploc=True).nodes()
    check_tree(nodes, n)


def _interleave(x, y, z):
    out = 0
    for b in range(21):
        out |= ((x >> b) & 1) << (3 * b)
        out |= ((y >> b) & 1) << (3 * b + 1)
        out |= ((z >> b) & 1) << (3 * b + 2)
    return out


def _as_u32(f):
    if not (f > 0):
        return 0
    return 0xFFFFFFFF if f >= 2.0 ** 32 else int(f)


def test_morton_codes():
    rng = np.random.default_rng(0)
    for x, y, z in list(rng.uniform(0, 1, (300, 3))) + [(1.0, 0.5, 0.0), (0.0, 0.0, 1.0), (-0.5, np.nan, 2.0)]:
        a, b, c = (_as_u32(v * 2.0 ** 21) for v in (x, y, z))
        assert orc.morton_encode_unorm(x, y, z) == _interleave(a & 0x1FFFFF, b & 0x1FFFFF, c & 0x1FFFFF)
    # a coordinate of exactly 1.0 reaches 2^21, which the 21-bit mask wraps to 0
    assert orc.morton_encode_unorm(1.0, 0.0, 0.0) == 0
    assert orc.morton_encode_unorm(0.5, 0.0, 0.0) == 1 << 60
    assert orc.morton_encode_unorm(0.0, 0.5, 0.0) == 1 << 61
    assert orc.morton_encode_unorm(0.0, 0.0, 0.5) == 1 << 62


def _py_ploc(aabbs):
    """Independent float32 implementation of from_leaves(Ploc) for the test."""
    f = np.float32
    n = aabbs.shape[0]
    mins, maxs = aabbs[:, :3].astype(f), aabbs[:, 3:].astype(f)
    cen = (mins + maxs) * f(0.5)
    cmin, cmax = cen.min(0), cen.max(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = f(1) / (cmax - cmin)
        rel = (cen - cmin) * inv
    keys = [_interleave(*(_as_u32(float(v) * 2.0 ** 21) & 0x1FFFFF for v in r)) for r in rel]
    order = sorted(range(n), key=lambda i: keys[i])  # Python's sort is stable
    # a node: (mins, maxs, children, leaf_count, is_leaf); wide nodes by creation id
    cur = [(mins[i], maxs[i], i, 1) for i in order]
    wide = {0: None}
    while len(cur) > 1:
        L = len(cur)
        cand = []
        for i in range(L):
            best, bk = f(3.4028235e38), -1
            for k in range(max(0, i - 16), min(L - 1, i + 16) + 1):
                if k == i:
                    continue
                e = np.maximum(cur[i][1], cur[k][1]) - np.minimum(cur[i][0], cur[k][0])
                sah = e[0] * (e[1] + e[2]) + e[1] * e[2]
                if sah < best:
                    best, bk = sah, k
            cand.append(bk)
        nxt = []
        for i in range(L):
            k = cand[i]
            if cand[k] == i:
                if i > k:
                    continue
                if L == 2:
                    wide[0] = (cur[i], cur[k])
                else:
                    wid = len(wide)
                    wide[wid] = (cur[i], cur[k])
                    nxt.append((np.minimum(cur[i][0], cur[k][0]), np.maximum(cur[i][1], cur[k][1]), wid,
                                cur[i][3] + cur[k][3]))
            else:
                nxt.append(cur[i])
        cur = nxt
    # depth-first relayout, left first
    out = []

    def visit(wid):
        slot = len(out)
        out.append(None)
        rec = []
        for node in wide[wid]:
            if node[3] == 1:
                rec.append((node[0], node[1], node[2], 1))
            else:
                rec.append((node[0], node[1], visit(node[2]), node[3]))
        out[slot] = rec
        return slot

    visit(0)
    return out


def test_tree_equals_independent_implementation():
    rng = np.random.default_rng(5)
    for n, spread in ((3, 1.0), (17, 4.0), (60, 10.0), (200, 30.0), (333, 2.0)):
        c = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
        h = rng.uniform(0.05, 1.0, (n, 3)).astype(np.float32)
        aabbs = np.concatenate([c - h, c + h], 1).astype(np.float32)
        if n == 333:  # repeated centres: equal keys keep the input order
            aabbs[100:150] = aabbs[100]
        got = orc.OracleBvh(aabbs, ploc=True).nodes()
        ref = _py_ploc(aabbs)
        check_tree(got, n)
        assert len(ref) == got.shape[0]
        for w, rec in enumerate(ref):
            for side, (mn, mx, ch, lc) in zip(("left", "right"), rec):
                assert np.array_equal(got[w][side + "_mins"], mn), (n, w, side)
                assert np.array_equal(got[w][side + "_maxs"], mx), (n, w, side)
                assert int(got[w][side + "_children"]) == ch, (n, w, side)
                assert int(got[w][side + "_data"]) & LC == lc, (n, w, side)
        # the PLOC tree is not the binned one (a different build really ran)
        if n >= 60:
            assert not np.array_equal(got, orc.OracleBvh(aabbs).nodes())


def test_ploc_tree_serves_self_pairs():
    rng = np.random.default_rng(9)
    n = 3000
    c = rng.uniform(0, 40, (n, 3)).astype(np.float32)
    h = rng.uniform(0.2, 1.5, (n, 3)).astype(np.float32)
    aabbs = np.concatenate([c - h, c + h], 1)
    b = orc.OracleBvh(aabbs, ploc=True)
    check_tree(b.nodes(), n)
    pairs, _ = b.self_pairs()
    assert orc.pair_set(pairs) == orc.pair_set(orc.brute_pairs(aabbs))
    assert len(pairs) > 1000
