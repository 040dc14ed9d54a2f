"""Trains the GJK warm-start network on the GPU and writes its parameters.

    python -m parry_amd.train_warmstart --pairs 8388608 --out gpurun_out/warmstart_mlp.npy

Training data: mixed convex scenes (``scenes.mixed_scene``: ConvexPolyhedron
with 8-64 hull points, Cuboid, Ball, Capsule, Cone, uniform types, centres in a
cube sized for ~8 AABB overlaps per shape), their broadphase pairs without a
Ball (the pairs GJK runs on with a seed), each labelled with the direction the
plain GJK produces (normal1 in shape 1's frame).  The network is trained with
the library's own MFMA forward/backward kernels and Adam.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import multiprocessing as mp
import os
import sys
import time

import numpy as np


def _scene(args):
    n, seed = args
    from parry_amd import scenes

    return scenes.mixed_scene(n, seed)


def build_training_set(num_pairs: int, shapes_per_scene: int, first_seed: int, workers: int, log=print):
    """(features, labels) CUDA tensors with at least ``num_pairs`` rows (then cut)."""
    import torch

    from parry_amd import Bvh, ShapeSet
    from parry_amd.learned import training_set

    feats, labels, total = [], [], 0
    seed = first_seed
    ctx = mp.get_context("spawn")  # scene generation is numpy/scipy only; no GPU in the workers
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as pool:
        pending = [pool.submit(_scene, (shapes_per_scene, seed + k)) for k in range(workers)]
        seed += workers
        while total < num_pairs:
            sc = pending.pop(0).result()
            pending.append(pool.submit(_scene, (shapes_per_scene, seed)))
            seed += 1
            shapes = ShapeSet(sc["descs"], sc["vertices"])
            poses = torch.from_numpy(sc["poses"]).cuda()
            aabbs = shapes.compute_aabbs(poses)
            pairs = Bvh(aabbs.cpu().numpy()).self_pairs()
            f, y = training_set(shapes, sc["descs"], poses, pairs)
            feats.append(f)
            labels.append(y)
            total += f.shape[0]
            log(f"scene {seed - workers - 1}: {pairs.shape[0]} pairs, {f.shape[0]} labelled, total {total}")
        for p in pending:
            p.cancel()
    X = torch.cat(feats)[:num_pairs].contiguous()
    Y = torch.cat(labels)[:num_pairs].contiguous()
    return X, Y


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=8 * 1024 * 1024)
    ap.add_argument("--shapes-per-scene", type=int, default=262144)
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=5e-5)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default="gpurun_out/warmstart_mlp.npy")
    args = ap.parse_args(argv)

    import torch

    from parry_amd.learned import WarmStartModel

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    X, Y = build_training_set(args.pairs, args.shapes_per_scene, args.first_seed, args.workers, log)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"training set: {X.shape[0]} pairs in {t1 - t0:.1f} s")
    net = WarmStartModel(seed=7)
    losses = net.train(X, Y, epochs=args.epochs, batch=args.batch, lr=args.lr, lr_final=args.lr_final, seed=11)
    t2 = time.perf_counter()
    log(f"trained {args.epochs} epochs in {t2 - t1:.1f} s; mean loss per epoch (1 - cos): {list(losses)}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, net.params())
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
