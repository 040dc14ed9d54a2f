"""Writes epa_shallow_pairs.npz: four overlapping polytope pairs of the dynamic
scene's first frame (mixed_scene(262144, seed=4), the oracle's Bvh pair order)
on which EPA's converged exit reports the best face's plane distance, shallower
than the exact depth, in f32 (pairs 101170, 402029) or in f64 (895187, 518705).

Inputs only: the two shapes' records, vertices and poses of each pair.
Run from the repository root: python tests/golden/make_epa_shallow_pairs.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from parry_amd import scenes  # noqa: E402

PAIRS = (101170, 402029, 895187, 518705)


def main():
    sc = scenes.mixed_scene(262144, seed=4)
    aabbs = oracle.shape_aabbs(sc["descs"], sc["vertices"], sc["poses"])
    pairs, _ = oracle.OracleBvh(aabbs).self_pairs()
    descs, verts, poses = [], [], []
    nv = 0
    for k in PAIRS:
        for i in pairs[k]:
            d = sc["descs"][i].copy()
            if d["type"] == 4:
                v = sc["vertices"][d["first_vertex"]:d["first_vertex"] + d["num_vertices"]]
                d["first_vertex"] = nv
                verts.append(v)
                nv += v.shape[0]
            descs.append(d)
            poses.append(sc["poses"][i])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "epa_shallow_pairs.npz")
    np.savez(out, descs=np.array(descs), vertices=np.concatenate(verts).astype(np.float32),
             poses=np.array(poses, np.float32), scene_pairs=np.array(PAIRS))
    print("wrote", out)


if __name__ == "__main__":
    main()
