"""The learned GJK warm start on the GPU, against the CPU oracle.

  * the bf16 MFMA forward kernel and the training gradients (forward, loss,
    backward, weight-gradient reduction) against oracle/mlp_oracle.py, which
    rounds to bf16 at the same points and accumulates in float64;
  * the fused seed kernel (features + network + normalisation) against the
    feature-row path; shape descriptors against their closed forms;
  * query::contact with network seeds against the oracle given the same seeds
    (same bar as every contact: has_contact identical, values within 1e-5);
  * training on plain-GJK directions lowers the loss and the GJK iteration count
    on held-out pairs, while the contacts stay those of the plain GJK within
    GJK's own tolerance.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from parry_amd import ShapeSet, scenes
from parry_amd.learned import (WarmStartModel, gjk_directions, pair_features, support_map_pairs,
                               training_set)

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def quat_matrix(q):
    x, y, z, w = q
    return np.array(
        [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _random_rows(rng, n, scale=1.0):
    X = np.zeros((n, 48), np.float32)
    X[:, :44] = rng.standard_normal((n, 44)) * scale
    Y = rng.standard_normal((n, 3))
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    return X, Y.astype(np.float32)


def test_forward_matches_oracle():
    torch = _torch()
    rng = np.random.default_rng(1)
    p = mo.he_init(2)
    net = WarmStartModel(p)
    assert np.array_equal(net.params(), p)
    for n in (1, 31, 32, 33, 1000, 70001):
        X, _ = _random_rows(rng, n)
        print(f"case: forward n={n}", flush=True)
        out = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
        ref, _ = mo.forward(p, X[:, :44], emulate=True)
        err = np.abs(out - ref).max() / max(1.0, np.abs(ref).max())
        assert err < 2e-3, (n, err)


def test_gradients_match_oracle():
    torch = _torch()
    rng = np.random.default_rng(3)
    p = mo.he_init(4)
    net = WarmStartModel(p)
    for n in (32, 100, 4096):
        X, Y = _random_rows(rng, n)
        rows = rng.permutation(n).astype(np.uint32)
        print(f"case: gradients n={n}", flush=True)
        loss_sum, g = net.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda(),
                                    torch.from_numpy(rows.view(np.int32)).cuda())
        g = g.cpu().numpy().astype(np.float64)
        lref, gref = mo.loss_and_grads(p, X[rows, :44], Y[rows], emulate=True)
        assert abs(loss_sum / n - lref) <= 1e-4 * max(1.0, abs(lref))
        rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
        big = np.abs(gref) > 1e-3 * np.abs(gref).max()
        elem = np.abs(g - gref)[big] / np.abs(gref)[big]
        print(f"n={n}: gradient rel. norm error {rel:.2e}, 99th pct element error {np.quantile(elem, 0.99):.2e}")
        assert rel < 5e-3 and np.quantile(elem, 0.99) < 2e-2


def _small_scene(n, seed):
    torch = _torch()
    sc = scenes.mixed_scene(n, seed)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    from parry_amd import Bvh

    aabbs = shapes.compute_aabbs(poses)
    pairs = Bvh(aabbs.cpu().numpy()).self_pairs()
    return sc, shapes, poses, pairs


def test_descriptors_and_fused_seeds():
    torch = _torch()
    sc, shapes, poses, pairs = _small_scene(3000, 7)
    d = shapes.descriptors().cpu().numpy()
    descs, verts = sc["descs"], sc["vertices"]
    t = descs["type"]
    onehot = {0: 0, 1: 1, 2: 2, 3: 3, 4: 4}
    for i in range(len(descs)):
        assert d[i, onehot[int(t[i])]] == 1.0 and d[i, :5].sum() == 1.0
        p = descs["params"][i]
        if t[i] == 4:
            v = verts[descs["first_vertex"][i]:descs["first_vertex"][i] + descs["num_vertices"][i]]
            assert np.allclose(d[i, 8:11], (v.max(0) - v.min(0)) / 2, atol=1e-6)
            assert np.allclose(d[i, 14], np.linalg.norm(v, axis=1).max(), atol=1e-6)
        elif t[i] == 1:
            assert np.allclose(d[i, 8:11], p[:3]) and np.isclose(d[i, 14], np.linalg.norm(p[:3]))
    net = WarmStartModel(seed=5)
    desc = shapes.descriptors()
    seeds = net.seeds(poses, pairs, desc).cpu().numpy()
    feats = pair_features(poses, pairs, desc)
    raw = net.forward(feats).cpu().numpy()
    nrm = np.linalg.norm(raw.astype(np.float64), axis=1, keepdims=True)
    assert np.allclose(seeds, raw / nrm, atol=1e-6)
    # feature rows against a numpy restatement
    f = feats.cpu().numpy()
    pr = pairs.cpu().numpy().view(np.uint32)
    P = sc["poses"].astype(np.float64)
    k = 17
    i, j = pr[k]
    Ri, Rj = quat_matrix(P[i, :4]), quat_matrix(P[j, :4])
    R12 = Ri.T @ Rj
    assert np.allclose(f[k, :9], R12.reshape(-1), atol=1e-5)
    assert np.allclose(f[k, 9:12], Ri.T @ (P[j, 4:] - P[i, 4:]), atol=1e-4)
    assert np.array_equal(f[k, 12:28], d[i]) and np.array_equal(f[k, 28:44], d[j])
    assert (f[:, 44:] == 0).all()


def _contacts(shapes, poses, pairs, seeds=None, with_iters=True):
    from parry_amd import contact_pairs

    out, iters = contact_pairs(shapes, poses, pairs, 0.01, init_dirs=seeds, with_iters=True)
    return out, iters


def test_seeded_contacts_match_oracle(orc):
    torch = _torch()
    sc, shapes, poses, pairs = _small_scene(4000, 9)
    net = WarmStartModel(seed=11)
    seeds = net.seeds(poses, pairs, shapes.descriptors())
    out, iters = _contacts(shapes, poses, pairs, seeds)
    torch.cuda.synchronize()
    pr = pairs.cpu().numpy().view(np.uint32)
    sd = seeds.cpu().numpy()
    print(f"case: seeded contacts, {pr.shape[0]} pairs", flush=True)
    ref = orc.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pr, 0.01, init_dirs=sd)
    o = out.cpu().numpy()
    has = o.view(np.uint32)[:, 13]
    assert np.array_equal(has.astype(np.int64), ref["status"].astype(np.int64))
    c = ref["status"] == 1
    err = np.abs(o[c, :13] - ref["out"][c]) / np.maximum(1.0, np.abs(ref["out"][c]))
    assert err.max(initial=0) <= 1e-5
    assert np.array_equal(iters.cpu().numpy()[:, 0], ref["iters"][:, 0].astype(np.int32))


def test_training_lowers_loss_and_gjk_iterations():
    """A short training run on one scene; evaluation on another."""
    torch = _torch()
    sc, shapes, poses, pairs = _small_scene(40000, 21)
    feats, labels = training_set(shapes, sc["descs"], poses, pairs)
    print(f"case: training on {feats.shape[0]} pairs", flush=True)
    net = WarmStartModel(seed=3)
    rows = torch.arange(min(8192, feats.shape[0]), dtype=torch.int32, device="cuda")
    untrained = net.gradients(feats, labels, rows)[0] / rows.shape[0]
    losses = net.train(feats, labels, epochs=4, batch=2048, lr=2e-3, lr_final=2e-4, seed=1)
    print(f"untrained loss {untrained:.4f}; epoch losses {losses}")
    # mean 1 - cos: about 1 for an untrained network
    assert losses[-1] < losses[0] and losses[-1] < 0.3 * untrained
    sc2, shapes2, poses2, pairs2 = _small_scene(40000, 22)
    keep = torch.from_numpy(support_map_pairs(sc2["descs"], pairs2.cpu().numpy().view(np.uint32))).cuda()
    sp = pairs2[keep].contiguous()
    plain, it_plain = _contacts(shapes2, poses2, sp)
    seeds = net.seeds(poses2, sp, shapes2.descriptors())
    learned, it_learned = _contacts(shapes2, poses2, sp, seeds)
    torch.cuda.synchronize()
    ip, il = it_plain[:, 0].float().mean().item(), it_learned[:, 0].float().mean().item()
    print(f"held-out: {sp.shape[0]} pairs, mean GJK iterations plain {ip:.3f} learned {il:.3f}")
    assert il < ip
    hp = plain.view(torch.int32)[:, 13].cpu().numpy()
    hl = learned.view(torch.int32)[:, 13].cpu().numpy()
    agree = (hp == hl).mean()
    both = (hp == 1) & (hl == 1)
    dp, dl = plain[:, 12].cpu().numpy()[both], learned[:, 12].cpu().numpy()[both]
    derr = np.abs(dp - dl) / np.maximum(1.0, np.abs(dp))
    print(f"has_contact agreement {agree:.6f}; dist rel. difference 99th pct {np.quantile(derr, 0.99):.2e} "
          f"max {derr.max():.2e}")
    assert agree > 0.999
