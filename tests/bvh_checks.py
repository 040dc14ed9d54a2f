"""Structural checks of exported Bvh trees, shared by the Bvh tests."""
import numpy as np


def check_tree(nodes, n_leaves):
    """Bvh::assert_well_formed (bvh_validation.rs:61-134, 196-218) on BVH_NODE_WIDE_DTYPE
    records: depth-first preorder ids, leaf counts, boxes that are the exact merge
    of the children's, every leaf exactly once."""
    lc = lambda d: d & 0x3FFFFFFF
    assert nodes.shape[0] == max(1, n_leaves - 1)
    seen = set()
    # depth-first order (bvh_validation.rs:196-218)
    stack, expect = [0], 0
    while stack:
        i = stack.pop()
        assert i == expect
        expect += 1
        n = nodes[i]
        # a slot with leaf count 0 is the empty right half of a one-leaf root
        if lc(n["right_data"]) > 1:
            stack.append(int(n["right_children"]))
        if lc(n["left_data"]) > 1:
            stack.append(int(n["left_children"]))

    # leaf totals bottom-up: in depth-first preorder (checked above) every child
    # has a larger id than its parent, so one reverse sweep sees children first
    total = np.zeros(nodes.shape[0], np.int64)
    for i in range(nodes.shape[0] - 1, -1, -1):
        n = nodes[i]
        t = 0
        for side in ("left", "right"):
            d = lc(int(n[side + "_data"]))
            c = int(n[side + "_children"])
            if d == 1:
                assert c not in seen
                seen.add(c)
                t += 1
            elif d > 1:
                ch = nodes[c]
                assert d == total[c]
                for cs in ("left", "right"):
                    assert (n[side + "_mins"] <= ch[cs + "_mins"]).all()
                    assert (n[side + "_maxs"] >= ch[cs + "_maxs"]).all()
                assert np.array_equal(n[side + "_mins"], np.minimum(ch["left_mins"], ch["right_mins"]))
                assert np.array_equal(n[side + "_maxs"], np.maximum(ch["left_maxs"], ch["right_maxs"]))
                t += d
        total[i] = t

    if n_leaves >= 2:
        assert total[0] == n_leaves
        assert seen == set(range(n_leaves))
    elif n_leaves == 1:
        assert lc(int(nodes[0]["left_data"])) == 1 and int(nodes[0]["left_children"]) == 0
        assert lc(int(nodes[0]["right_data"])) == 0
