"""Seeded synthetic workloads for the collision-query hot path.

These generators build the inputs that the benchmarks and the parity tests
share: noise-displaced icospheres with ray batches aimed into them, and (in
later sections) convex shape populations.  Everything is numpy, seeded with
``numpy.random.Generator(PCG64(seed))``, computed in float64 and rounded to
float32 once at the end, so the same seed always gives bit-identical float32
inputs on every machine.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "icosahedron",
    "icosphere",
    "noisy_icosphere",
    "rays_into_unit_ball",
    "load_obj",
]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def icosahedron() -> tuple[np.ndarray, np.ndarray]:
    """Unit icosahedron (12 vertices, 20 faces) with outward counter-clockwise winding."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [
            [-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0],
            [0, -1, phi], [0, 1, phi], [0, -1, -phi], [0, 1, -phi],
            [phi, 0, -1], [phi, 0, 1], [-phi, 0, -1], [-phi, 0, 1],
        ],
        dtype=np.float64,
    )
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    # Orient every face outward (normal . centroid > 0).
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    c = v[f].mean(axis=1)
    flip = (n * c).sum(axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: midpoint subdivision with shared-edge deduplication.

    Level s has 10*4^s + 2 vertices and 20*4^s triangles (s=6: 40,962 / 81,920).
    Each triangle (a, b, c) becomes (a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca),
    so the outward winding is preserved.  Returns float64 vertices on the unit
    sphere and int64 indices.
    """
    v, f = icosahedron()
    for _ in range(subdivisions):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key[:, 0] * (nv + 1) + key[:, 1], return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        mid_idx = inv.reshape(3, -1).T + nv  # columns: ab, bc, ca
        ab, bc, ca = mid_idx[:, 0], mid_idx[:, 1], mid_idx[:, 2]
        A, B, C = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate(
            [
                np.stack([A, ab, ca], axis=1),
                np.stack([B, bc, ab], axis=1),
                np.stack([C, ca, bc], axis=1),
                np.stack([ab, bc, ca], axis=1),
            ],
            axis=0,
        )
        v = np.concatenate([v, mid], axis=0)
    return v, f


def noisy_icosphere(subdivisions: int, seed: int, radius: float = 1.0, noise: float = 0.05):
    """Icosphere of the given level, each vertex scaled radially by 1 + U[-noise, noise].

    Returns (vertices float32 [nv,3], indices uint32 [nt,3]).
    """
    v, f = icosphere(subdivisions)
    r = _rng(seed).uniform(-noise, noise, size=v.shape[0])
    v = v * (radius * (1.0 + r))[:, None]
    return np.ascontiguousarray(v, dtype=np.float32), np.ascontiguousarray(f, dtype=np.uint32)


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    g = rng.standard_normal(size=(n, 3))
    nrm = np.linalg.norm(g, axis=1, keepdims=True)
    nrm[nrm == 0] = 1.0
    return g / nrm


def rays_into_unit_ball(n: int, seed: int, origin_radius: float = 3.0) -> np.ndarray:
    """Rays with origins uniform on a sphere of radius ``origin_radius`` aimed at
    points uniform in the unit ball; directions are normalised.

    Returns float32 [n, 6] laid out as parry's ``Ray`` (origin xyz, dir xyz).
    """
    rng = _rng(seed)
    o = _unit_vectors(rng, n) * origin_radius
    t = _unit_vectors(rng, n) * (rng.uniform(0.0, 1.0, size=(n, 1)) ** (1.0 / 3.0))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([o, d], axis=1), dtype=np.float32)


def uneven_mesh_batch(num_meshes: int, seed: int, levels=(3, 7), log2_rays=(10.0, 20.0)):
    """The uneven multi-mesh ray workload: per mesh a subdivision level uniform
    over ``levels`` (inclusive) and 2^X rays with X uniform in ``log2_rays``
    (continuous, rounded to the nearest integer count).  64 meshes at the
    default ranges give about 10 M rays over about 5.6 M triangles.

    Returns a list of (level, num_rays, mesh_seed, ray_seed) tuples; build the
    meshes with ``noisy_icosphere(level, mesh_seed)`` and the rays with
    ``rays_into_unit_ball(num_rays, ray_seed)``.
    """
    rng = _rng(seed)
    lv = rng.integers(levels[0], levels[1] + 1, size=num_meshes)
    x = rng.uniform(log2_rays[0], log2_rays[1], size=num_meshes)
    counts = np.rint(np.exp2(x)).astype(np.int64)
    return [(int(lv[i]), int(counts[i]), seed * 1000 + 2 * i + 1, seed * 1000 + 2 * i + 2) for i in range(num_meshes)]


def load_obj(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Minimal Wavefront OBJ reader (``v`` and triangular ``f`` records, 1-based)."""
    vs, fs = [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                fs.append([int(x.split("/")[0]) - 1 for x in p[1:4]])
    return np.asarray(vs, dtype=np.float32), np.asarray(fs, dtype=np.uint32)


# ---------------------------------------------------------------------------
# Convex shape populations (broadphase + contact workloads)
# ---------------------------------------------------------------------------

def random_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's method), float64 [n, 4] as (i, j, k, w)."""
    u1, u2, u3 = rng.uniform(0.0, 1.0, size=(3, n))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack(
        [a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)],
        axis=1,
    )
    return q


def convex_point_cloud(rng: np.random.Generator, k: int) -> np.ndarray:
    """k points on an ellipsoid (semi-axes U[0.5, 1.5]) jittered radially by U[-5%, +5%]."""
    axes = rng.uniform(0.5, 1.5, size=3)
    d = _unit_vectors(rng, k)
    return d * axes * (1.0 + rng.uniform(-0.05, 0.05, size=(k, 1)))


def remove_unused_points_order(num_points: int, used) -> list[int]:
    """The input indices ``remove_unused_points`` (src/utils/cleanup.rs:6-34) keeps,
    in the order it leaves them: scanning up, each unused point is replaced by
    the current last point (``Vec::swap_remove``) and its slot is examined again.
    This is the vertex order of ``ConvexPolyhedron::from_convex_hull``
    (``try_convex_hull``, convex_hull.rs:238-242, keeps the caller's points and
    drops the unused ones this way; ``from_convex_mesh`` keeps that order)."""
    order = list(range(num_points))
    flags = [bool(u) for u in used]
    i = 0
    while i != len(order):
        if not flags[i]:
            last = order.pop()
            if i < len(order):
                order[i] = last
            flags[i] = flags[len(order)]
        else:
            i += 1
    return order


def convex_hull_vertices(pts: np.ndarray) -> np.ndarray:
    """Vertices of the convex hull of ``pts``, in the order parry's
    ``ConvexPolyhedron::from_convex_hull`` stores them (``remove_unused_points_order``).

    The vertex set is scipy's Qhull (float64).  parry's own hull (f32, points
    scaled by their AABB diagonal) treats a point as outside a facet only beyond
    100 EPS of it (triangle_facet.rs:83-97), so a hull vertex that sticks out of
    the hull of the others by less than about 1.2e-5 of the diagonal, or an
    interior point that close to the surface, can be classified differently:
    about 1% of the 8-64-point clouds of ``convex_point_cloud`` hold such a point."""
    from scipy.spatial import ConvexHull

    h = ConvexHull(pts)
    used = np.zeros(pts.shape[0], bool)
    used[h.vertices] = True
    return pts[remove_unused_points_order(pts.shape[0], used)]


def convex_polyhedra(n: int, seed: int, kmin: int = 8, kmax: int = 64):
    """``n`` ConvexPolyhedron vertex sets: the hull of kmin..kmax (uniform) points
    sampled as in ``convex_point_cloud``.  Returns (vertices float32 [V, 3], offsets uint32 [n+1])."""
    rng = _rng(seed)
    ks = rng.integers(kmin, kmax + 1, size=n)
    verts, offs = [], [0]
    for k in ks:
        v = convex_hull_vertices(convex_point_cloud(rng, int(k)))
        verts.append(v)
        offs.append(offs[-1] + v.shape[0])
    return np.ascontiguousarray(np.concatenate(verts, 0), dtype=np.float32), np.asarray(offs, dtype=np.uint32)


def _rotated_half_extents(verts: np.ndarray, offs: np.ndarray, quats: np.ndarray) -> np.ndarray:
    """Half extents of each shape's AABB under its rotation (float64, for calibration only)."""
    n = offs.shape[0] - 1
    out = np.zeros((n, 3))
    for i in range(n):
        p = verts[offs[i]:offs[i + 1]].astype(np.float64)
        x, y, z, w = quats[i]
        R = np.array(
            [
                [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
            ]
        )
        q = p @ R.T
        out[i] = (q.max(0) - q.min(0)) / 2
    return out


def calibrate_cube_side(half_extents: np.ndarray, target_overlaps: float, rng: np.random.Generator,
                        samples: int = 200000) -> float:
    """Side L of the cube in which uniform centres give ``target_overlaps`` expected AABB
    overlaps per shape: for two centres uniform on [0, L] and summed half-width s,
    P(|x_i - x_j| <= s) = 2s/L - s^2/L^2; the expectation over shape pairs is
    estimated from ``samples`` random pairs and L is found by bisection."""
    n = half_extents.shape[0]
    i = rng.integers(0, n, samples)
    j = rng.integers(0, n, samples)
    s = half_extents[i] + half_extents[j]

    def overlaps(L):
        p = np.clip(2 * s / L - (s / L) ** 2, 0.0, 1.0)
        return (n - 1) * p.prod(axis=1).mean()

    lo, hi = float(s.max()), 1e6
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if overlaps(mid) > target_overlaps:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def polyhedra_scene(n: int, seed: int, target_overlaps: float = 8.0):
    """The convex-polyhedra broadphase workload: ``n`` ConvexPolyhedra (8-64 hull
    points on jittered ellipsoids), uniform random rotations and centres uniform
    in a cube calibrated for ``target_overlaps`` mean AABB overlaps per shape.

    Returns dict(vertices float32 [V,3], offsets uint32 [n+1], poses float32 [n,7]
    as parry Isometry3 (i, j, k, w, tx, ty, tz), side float).
    """
    verts, offs = convex_polyhedra(n, seed)
    rng = _rng(seed + 1)
    q = random_rotations(n, rng)
    he = _rotated_half_extents(verts, offs, q)
    L = calibrate_cube_side(he, target_overlaps, rng)
    t = rng.uniform(0.0, L, size=(n, 3))
    poses = np.ascontiguousarray(np.concatenate([q, t], 1), dtype=np.float32)
    return {"vertices": verts, "offsets": offs, "poses": poses, "side": L}


MIXED_TYPES = ("convex_polyhedron", "cuboid", "ball", "capsule", "cone")


def mixed_shapes(n: int, seed: int, kmin: int = 8, kmax: int = 64):
    """``n`` shapes with types uniform over ConvexPolyhedron (kmin..kmax hull
    points on a jittered ellipsoid, semi-axes U[0.5, 1.5]), Cuboid (half extents
    U[0.5, 1.5]^3), Ball (radius U[0.5, 1.5]), Capsule (half height U[0.5, 1.5],
    radius U[0.25, 0.75]) and Cone (half height and radius U[0.5, 1.5]).

    Returns (descs SHAPE_DESC_DTYPE [n], vertices float32 [V, 3]).
    """
    from .shape import SHAPE_DESC_DTYPE

    rng = _rng(seed)
    kinds = rng.integers(0, 5, size=n)
    descs = np.zeros(n, dtype=SHAPE_DESC_DTYPE)
    verts = []
    nv = 0
    type_code = {0: 4, 1: 1, 2: 0, 3: 2, 4: 3}  # -> PQ_SHAPE_* codes
    for i, k in enumerate(kinds):
        descs[i]["type"] = type_code[int(k)]
        if k == 0:
            v = convex_hull_vertices(convex_point_cloud(rng, int(rng.integers(kmin, kmax + 1))))
            descs[i]["first_vertex"] = nv
            descs[i]["num_vertices"] = v.shape[0]
            verts.append(v)
            nv += v.shape[0]
        elif k == 1:
            descs[i]["params"][:3] = rng.uniform(0.5, 1.5, 3)
        elif k == 2:
            descs[i]["params"][0] = rng.uniform(0.5, 1.5)
        elif k == 3:
            descs[i]["params"][:2] = [rng.uniform(0.5, 1.5), rng.uniform(0.25, 0.75)]
        else:
            descs[i]["params"][:2] = rng.uniform(0.5, 1.5, 2)
    v = np.concatenate(verts, 0) if verts else np.zeros((0, 3))
    return descs, np.ascontiguousarray(v, dtype=np.float32)


def _mixed_half_extents(descs, verts, quats):
    """World AABB half extents of each mixed shape under its rotation (float64,
    for the cube calibration only)."""
    q = np.asarray(quats, np.float64)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)  # (n, 3, 3)
    t = descs["type"]
    p = descs["params"].astype(np.float64)
    he = np.zeros((len(descs), 3))
    # Cuboid: |R| he; Ball: r; Capsule: |R e_y| h + r; Cone: |R e_y| h + r sqrt(1 - (R e_y)^2)
    cub = t == 1
    he[cub] = np.einsum("nij,nj->ni", np.abs(R[cub]), p[cub, :3])
    ball = t == 0
    he[ball] = p[ball, :1]
    cap = t == 2
    he[cap] = np.abs(R[cap, :, 1]) * p[cap, :1] + p[cap, 1:2]
    cone = t == 3
    ay = R[cone, :, 1]
    he[cone] = np.abs(ay) * p[cone, :1] + p[cone, 1:2] * np.sqrt(np.clip(1 - ay * ay, 0, 1))
    poly = np.nonzero(t == 4)[0]
    if poly.size:
        first = descs["first_vertex"][poly].astype(np.int64)
        cnt = descs["num_vertices"][poly].astype(np.int64)
        owner = np.repeat(np.arange(poly.size), cnt)
        idx = np.concatenate([np.arange(f, f + c) for f, c in zip(first, cnt)])
        v = verts[idx].astype(np.float64)
        wv = np.einsum("nij,nj->ni", R[poly][owner], v)
        starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        mx = np.maximum.reduceat(wv, starts, axis=0)
        mn = np.minimum.reduceat(wv, starts, axis=0)
        he[poly] = (mx - mn) / 2
    return he


def mixed_scene(n: int, seed: int, target_overlaps: float = 8.0):
    """``n`` mixed convex shapes (``mixed_shapes``) with uniform random rotations
    and centres uniform in a cube calibrated for ``target_overlaps`` mean AABB
    overlaps per shape (the learned warm-start and dynamic-scene workloads).

    Returns dict(descs, vertices, poses float32 [n, 7] (i j k w tx ty tz), side).
    """
    descs, verts = mixed_shapes(n, seed)
    rng = _rng(seed + 1)
    q = random_rotations(n, rng)
    he = _mixed_half_extents(descs, verts, q)
    L = calibrate_cube_side(he, target_overlaps, rng)
    t = rng.uniform(0.0, L, size=(n, 3))
    poses = np.ascontiguousarray(np.concatenate([q, t], 1), dtype=np.float32)
    return {"descs": descs, "vertices": verts, "poses": poses, "side": L}



def shapes_around_mesh(n: int, seed: int, mesh_radius: float = 2.0, scale: float = 0.4,
                       shell=(0.3, 0.8)):
    """``mixed_shapes(n, seed)`` scaled by ``scale`` and placed around a mesh of
    radius ``mesh_radius`` centred at the origin: random rotations, centres at
    distance U[shell] x 2 mesh_radius along uniform random directions, so a part
    of the shapes straddles the surface and a part stands clear of it.

    Returns (descs, vertices float32 [V, 3], poses float32 [n, 7]).
    """
    descs, verts = mixed_shapes(n, seed)
    rng = _rng(seed + 1)
    q = random_rotations(n, rng)
    t = _unit_vectors(rng, n) * rng.uniform(shell[0], shell[1], (n, 1)) * (2.0 * mesh_radius)
    descs["params"] *= scale
    return descs, (verts * scale).astype(np.float32), np.concatenate([q, t], 1).astype(np.float32)


def approach_pairs(n: int, seed: int, spread=(2.0, 6.0), scale: float = 1.0, with_triangles: bool = True):
    """``n`` shape-cast pairs (2q, 2q+1) of ``mixed_shapes(2n, seed)``, every fifth
    shape replaced by a Triangle (vertices N(0, 0.7 scale)) when ``with_triangles``:
    shape 2q at a random rotation and a position uniform in [-50, 50]^3, shape 2q+1
    at a random rotation and U[spread] away along a uniform direction, moving back
    along it (with N(0, 0.3) noise) at U[0.5, 3] units per second; every shape
    also drifts at U[-0.3, 0.3]^3.

    Returns (descs, vertices float32, poses float32 [2n, 7], vels float32 [2n, 3],
    pairs uint32 [n, 2]).
    """
    rng = np.random.default_rng(seed)
    descs, verts = mixed_shapes(2 * n, seed)
    descs["params"] *= scale
    verts = verts * scale
    if with_triangles:  # every fifth shape becomes a Triangle
        t = np.arange(0, 2 * n, 5)
        tv = rng.normal(0.0, 0.7 * scale, (len(t), 3, 3)).astype(np.float32)
        descs["type"][t] = 5
        descs["first_vertex"][t] = verts.shape[0] + 3 * np.arange(len(t))
        descs["num_vertices"][t] = 3
        verts = np.concatenate([verts, tv.reshape(-1, 3)], 0)
    poses = np.concatenate([random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    dirs = rng.normal(0, 1, (n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    poses[1::2, 4:] = dirs * rng.uniform(spread[0], spread[1], (n, 1))
    poses[0::2, 4:] = rng.uniform(-50, 50, (n, 3))
    poses[1::2, 4:] += poses[0::2, 4:]
    vels = rng.uniform(-0.3, 0.3, (2 * n, 3))
    vels[1::2] += (-dirs + rng.normal(0, 0.3, (n, 3))) * rng.uniform(0.5, 3.0, (n, 1))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    return descs, verts.astype(np.float32), poses.astype(np.float32), vels.astype(np.float32), pairs


def feature_map_pairs(n: int, seed: int, spread: float = 0.6, types=(1, 2, 3, 5)):
    """``n`` contact-manifold pairs (2q, 2q+1) of shapes with polygonal feature maps:
    types drawn uniformly from ``types`` (Cuboid half extents U[0.3, 1]^3; Capsule
    and Cone half height and radius U[0.3, 1]; Triangle vertices N(0, 0.7); a Ball
    radius U[0.3, 1] when 0 is among the types), random rotations, shape 2q
    uniform in [-40, 40]^3 and shape 2q+1 offset from it by N(0, spread^2) per axis.

    Returns (descs, vertices float32, poses float32 [2n, 7], pairs uint32 [n, 2]).
    """
    from .shape import SHAPE_DESC_DTYPE

    rng = np.random.default_rng(seed)
    t = rng.choice(types, size=2 * n)
    d = np.zeros(2 * n, SHAPE_DESC_DTYPE)
    d["type"] = t
    cub = t == 1
    d["params"][cub, :3] = rng.uniform(0.3, 1.0, (int(cub.sum()), 3))
    cc = (t == 2) | (t == 3)
    d["params"][cc, :2] = rng.uniform(0.3, 1.0, (int(cc.sum()), 2))
    ball = t == 0
    d["params"][ball, 0] = rng.uniform(0.3, 1.0, int(ball.sum()))
    tri = np.nonzero(t == 5)[0]
    d["first_vertex"][tri] = 3 * np.arange(len(tri))
    d["num_vertices"][tri] = 3
    verts = rng.normal(0, 0.7, (3 * len(tri), 3)).astype(np.float32)
    poses = np.concatenate([random_rotations(2 * n, rng), np.zeros((2 * n, 3))], 1)
    poses[0::2, 4:] = rng.uniform(-40, 40, (n, 3))
    poses[1::2, 4:] = poses[0::2, 4:] + rng.normal(0, spread, (n, 3))
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    return d, verts, poses.astype(np.float32), pairs


def polyhedron_faces(descs, vertices, merge_tol: float = 1e-6):
    """Face tables of every ConvexPolyhedron in ``descs`` (others get no faces), in
    the layout ``ConvexPolyhedron`` keeps them (convex_polyhedron.rs:172-185, 395-540):
    one face per plane of the hull (coplanar hull triangles merged), its outward
    unit normal, its vertices counter-clockwise seen from outside, and the edge
    from its k-th to its (k+1)-th vertex at position k of ``face_edges``.  Edge ids
    are per polyhedron, numbered as faces first meet them.

    Returns the dict ``ShapeSet.set_polyhedron_faces`` takes.
    """
    from scipy.spatial import ConvexHull

    from .shape import POLYHEDRON_FACE_DTYPE, SHAPE_CONVEX_POLYHEDRON

    d = np.asarray(descs)
    v = np.asarray(vertices, np.float64).reshape(-1, 3)
    faces, fverts, fedges = [], [], []
    shape_faces = np.zeros((d.shape[0], 2), np.uint32)
    for i in np.nonzero(d["type"] == SHAPE_CONVEX_POLYHEDRON)[0]:
        first, nv = int(d["first_vertex"][i]), int(d["num_vertices"][i])
        pts = v[first:first + nv]
        h = ConvexHull(pts)
        # coplanar triangles of one hull face: same plane up to merge_tol
        key = np.round(h.equations / merge_tol).astype(np.int64)
        _, first_of, group = np.unique(key, axis=0, return_index=True, return_inverse=True)
        group = group.reshape(-1)
        shape_faces[i] = (len(faces), len(first_of))
        edge_ids: dict = {}
        for g in np.argsort(first_of, kind="stable"):  # faces in the order Qhull lists their planes
            tri = h.simplices[group == g]
            ids = np.unique(tri)
            nrm = h.equations[first_of[g], :3]
            c = pts[ids].mean(0)
            a = pts[ids[0]] - c
            a /= np.linalg.norm(a)
            b = np.cross(nrm, a)
            r = pts[ids] - c
            ids = ids[np.argsort(np.arctan2(r @ b, r @ a), kind="stable")]
            faces.append((nrm, len(fverts), len(ids)))
            for k in range(len(ids)):
                e = (min(ids[k], ids[(k + 1) % len(ids)]), max(ids[k], ids[(k + 1) % len(ids)]))
                fverts.append(int(ids[k]))
                fedges.append(edge_ids.setdefault(e, len(edge_ids)))
    f = np.zeros(len(faces), POLYHEDRON_FACE_DTYPE)
    for k, (nrm, fi, cnt) in enumerate(faces):
        f[k] = (nrm.astype(np.float32), fi, cnt)
    return {"faces": f, "face_vertices": np.asarray(fverts, np.uint32), "face_edges": np.asarray(fedges, np.uint32),
            "shape_faces": shape_faces}
