// contact_gjk.h -- device-side support maps, simplex projections and GJK (gfx950).
//
// These functions evaluate parry3d's f32 algorithms in parry's operation order
// (the file must be compiled with -ffp-contract=off):
//   * support maps: ConvexPolyhedron (first strict maximum of p.dir,
//     src/utils/point_cloud_support_point.rs:5-19), Ball (ball.rs:242-249),
//     Cuboid (cuboid.rs:454-456 + utils/wops.rs:30-35), Capsule
//     (capsule.rs:407-418), Cone (cone.rs:162-182), and the posed support
//     point of shape 2 (support_map.rs:380-383);
//   * CSOPoint::from_shapes (src/query/gjk/cso_point.rs:69-78);
//   * closest points of the origin on a segment / triangle / tetrahedron with
//     their Voronoi location (src/query/point/point_segment.rs:45-85,
//     point_triangle.rs:54-291, point_tetrahedron.rs:28-340);
//   * the 3D Voronoi simplex (src/query/gjk/voronoi_simplex3.rs:13-351);
//   * gjk::closest_points with exact_dist = true (src/query/gjk/gjk.rs:353-453)
//     and its result() (gjk.rs:797-818).
// One lane evaluates one pair; all state lives in registers.
#pragma once

#include "pq_internal.h"
#include "pq_math.h"

namespace pq {

constexpr float kEps = 1.1920929e-07f;      // f32::EPSILON = DEFAULT_EPSILON (src/lib.rs:102)
constexpr float kEpsTol = kEps * 10.0f;     // gjk::eps_tol (gjk.rs:141-144)
constexpr float kFltMax = 3.40282347e+38f;  // f32::MAX
constexpr uint32_t kGjkMaxIters = 100;      // gjk.rs:449-451
// parry's ConstantOrigin (src/query/gjk/special_support_maps.rs:32-54): the
// second operand of GJK/EPA point projections.  Internal, never in a shape set.
constexpr uint32_t kShapeOrigin = 6u;
// A Capsule's segment (0, -h, 0)-(0, h, 0): the feature map contact manifolds use
// for a Capsule (shape.rs:855-857).  Internal, never in a shape set.
constexpr uint32_t kShapeSegment = 7u;

struct ShapeRef {
    uint32_t type;
    uint32_t n;
    const float4 *v;
    float p0, p1, p2;
};

__device__ __forceinline__ ShapeRef load_shape(const ShapeRecord *rec, const float4 *verts, uint32_t i) {
    const ShapeRecord r = rec[i];
    return ShapeRef{r.type, r.num_vertices, verts + r.first_vertex, r.param[0], r.param[1], r.param[2]};
}

// One step of the polyhedron support scan: vertex p replaces the best only if
// strictly better (the first strict maximum wins, as in parry).
__device__ __forceinline__ void support_step(const float4 p, f3 d, f3 &best, float &best_dot) {
    const f3 c = f3{p.x, p.y, p.z};
    const float dd = dot3(c, d);
    const bool gt = dd > best_dot;
    best_dot = gt ? dd : best_dot;
    best.x = gt ? c.x : best.x;
    best.y = gt ? c.y : best.y;
    best.z = gt ? c.z : best.z;
}

__device__ __forceinline__ f3 local_support(const ShapeRef &s, f3 d) {
    if (s.type == PQ_SHAPE_CONVEX_POLYHEDRON) {
        const float4 p0 = s.v[0];
        f3 best = f3{p0.x, p0.y, p0.z};
        float best_dot = dot3(best, d);
        // Eight vertex loads in flight per step, compared in index order: the
        // scan's result is the sequential one, its latency an eighth of it.
        uint32_t k = 1;
        for (; k + 8 <= s.n; k += 8) {
            float4 q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = s.v[k + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) support_step(q[u], d, best, best_dot);
        }
        for (; k + 4 <= s.n; k += 4) {
            const float4 a = s.v[k], b = s.v[k + 1], c = s.v[k + 2], e = s.v[k + 3];
            support_step(a, d, best, best_dot);
            support_step(b, d, best, best_dot);
            support_step(c, d, best, best_dot);
            support_step(e, d, best, best_dot);
        }
        for (; k < s.n; ++k) support_step(s.v[k], d, best, best_dot);
        return best;
    }
    if (s.type == PQ_SHAPE_BALL) {
        const float l = sqrtf(norm_sq3(d));
        return mul3(div3(d, l), s.p0);
    }
    if (s.type == PQ_SHAPE_CUBOID) return f3{copysignf(s.p0, d.x), copysignf(s.p1, d.y), copysignf(s.p2, d.z)};
    if (s.type == PQ_SHAPE_TRIANGLE) {  // triangle.rs:697-715
        const float4 a = s.v[0], b = s.v[1], c = s.v[2];
        const float d1 = dot3(f3{a.x, a.y, a.z}, d), d2 = dot3(f3{b.x, b.y, b.z}, d), d3 = dot3(f3{c.x, c.y, c.z}, d);
        const float4 w = d1 > d2 ? (d1 > d3 ? a : c) : (d2 > d3 ? b : c);
        return f3{w.x, w.y, w.z};
    }
    if (s.type == kShapeOrigin) return f3{0.0f, 0.0f, 0.0f};
    if (s.type == kShapeSegment) {  // segment.rs:461-467
        const f3 a{0.0f, -s.p0, 0.0f}, b{0.0f, s.p0, 0.0f};
        return dot3(a, d) > dot3(b, d) ? a : b;
    }
    if (s.type == PQ_SHAPE_CAPSULE) {
        const float sq = norm_sq3(d);
        f3 u = sq > 0.0f ? div3(d, sqrtf(sq)) : f3{0.0f, 1.0f, 0.0f};
        const f3 a = f3{0.0f, -s.p0, 0.0f}, b = f3{0.0f, s.p0, 0.0f};
        return dot3(u, a) > dot3(u, b) ? add3(a, mul3(u, s.p1)) : add3(b, mul3(u, s.p1));
    }
    // PQ_SHAPE_CONE
    f3 v = f3{d.x, 0.0f, d.z};
    const float nrm = sqrtf(norm_sq3(v));
    v = div3(v, nrm);
    if (nrm == 0.0f) return f3{0.0f, copysignf(s.p0, d.y), 0.0f};
    v = mul3(v, s.p1);
    v.y = -s.p0;
    if (dot3(d, v) < d.y * s.p0) return f3{0.0f, s.p0, 0.0f};
    return v;
}

struct Cso {
    f3 p, o1, o2;
};

// CSOPoint::from_shapes: s1(dir) - pos12 * s2(pos12^-1 * -dir); ConstantOrigin's
// support_point is pos12's translation (special_support_maps.rs:36-38).
__device__ __forceinline__ Cso cso_from_shapes(const iso &pos12, const ShapeRef &g1, const ShapeRef &g2, f3 dir) {
    const f3 a = local_support(g1, dir);
    const f3 b = g2.type == kShapeOrigin ? pos12.t : iso_point(pos12, local_support(g2, qrot_inv(pos12.q, neg3(dir))));
    return Cso{sub3(a, b), a, b};
}

// Shape::compute_aabb (src/shape/shape.rs:369) of record r at pose m.
// ConvexPolyhedron: point_cloud_aabb (aabb_utils.rs:66-86) -- every vertex
// posed and folded into a running min/max from vertex 0.  Triangle: the posed
// vertices, a.min(b).min(c) (aabb_triangle.rs:10-30).  Ball: centre -/+ radius
// (aabb_ball.rs:7-27).  Cuboid: centre -/+ |R| * half_extents (aabb_cuboid.rs:9-14,
// utils/isometry_ops.rs:14-19) with R nalgebra's quaternion-to-matrix.  Capsule:
// the posed segment's AABB grown by the radius (aabb_capsule.rs:8-20).  Cone:
// support points along -/+ each axis (aabb_support_map.rs:12-14 -> aabb_utils.rs:10-31).
__device__ __forceinline__ void shape_aabb(const ShapeRecord &r, const float4 *verts, const iso &m, f3 &lo, f3 &hi) {
    if (r.type == PQ_SHAPE_CONVEX_POLYHEDRON) {
        const float4 *v = verts + r.first_vertex;
        float4 p0 = v[0];
        lo = hi = iso_point(m, f3{p0.x, p0.y, p0.z});
        for (uint32_t k = 1; k < r.num_vertices; ++k) {
            const float4 p = v[k];
            const f3 w = iso_point(m, f3{p.x, p.y, p.z});
            lo = min3(lo, w);
            hi = max3(hi, w);
        }
    } else if (r.type == PQ_SHAPE_TRIANGLE) {
        const float4 *v = verts + r.first_vertex;
        const f3 a = iso_point(m, f3{v[0].x, v[0].y, v[0].z});
        const f3 b = iso_point(m, f3{v[1].x, v[1].y, v[1].z});
        const f3 c = iso_point(m, f3{v[2].x, v[2].y, v[2].z});
        lo = min3(min3(a, b), c);
        hi = max3(max3(a, b), c);
    } else if (r.type == PQ_SHAPE_BALL) {
        const float rad = r.param[0];
        lo = add3(m.t, f3{-rad, -rad, -rad});
        hi = add3(m.t, f3{rad, rad, rad});
    } else if (r.type == PQ_SHAPE_CUBOID) {
        // UnitQuaternion::to_rotation_matrix, then abs(R) * half_extents.
        const float qi = m.q.x, qj = m.q.y, qk = m.q.z, qw = m.q.w;
        const float ww = qw * qw, ii = qi * qi, jj = qj * qj, kk = qk * qk;
        const float ij = qi * qj * 2.0f, wk = qw * qk * 2.0f, wj = qw * qj * 2.0f;
        const float ik = qi * qk * 2.0f, jk = qj * qk * 2.0f, wi = qw * qi * 2.0f;
        const float R[3][3] = {{((ww + ii) - jj) - kk, ij - wk, wj + ik},
                               {wk + ij, ((ww - ii) + jj) - kk, jk - wi},
                               {ik - wj, wi + jk, ((ww - ii) - jj) + kk}};
        const float he[3] = {r.param[0], r.param[1], r.param[2]};
        float w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = (fabsf(R[a][0]) * he[0] + fabsf(R[a][1]) * he[1]) + fabsf(R[a][2]) * he[2];
        lo = sub3(m.t, f3{w[0], w[1], w[2]});
        hi = add3(m.t, f3{w[0], w[1], w[2]});
    } else if (r.type == PQ_SHAPE_CAPSULE) {
        const f3 a = iso_point(m, f3{0.0f, -r.param[0], 0.0f});
        const f3 b = iso_point(m, f3{0.0f, r.param[0], 0.0f});
        const float rad = r.param[1];
        lo = sub3(min3(a, b), f3{rad, rad, rad});
        hi = add3(max3(a, b), f3{rad, rad, rad});
    } else {  // PQ_SHAPE_CONE
        const ShapeRef sr{r.type, 0, nullptr, r.param[0], r.param[1], r.param[2]};
        float mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            f3 basis = f3{0.0f, 0.0f, 0.0f};
            if (d == 0) basis.x = 1.0f;
            if (d == 1) basis.y = 1.0f;
            if (d == 2) basis.z = 1.0f;
            const f3 pmax = iso_point(m, local_support(sr, qrot_inv(m.q, basis)));
            const f3 nb = f3{-basis.x, -basis.y, -basis.z};
            const f3 pmin = iso_point(m, local_support(sr, qrot_inv(m.q, nb)));
            mx[d] = d == 0 ? pmax.x : (d == 1 ? pmax.y : pmax.z);
            mn[d] = d == 0 ? pmin.x : (d == 1 ? pmin.y : pmin.z);
        }
        lo = f3{mn[0], mn[1], mn[2]};
        hi = f3{mx[0], mx[1], mx[2]};
    }
}

// relative_eq!(proj, origin) with approx's default tolerances.
__device__ __forceinline__ bool near_origin(f3 p) {
    return ((unsigned)(fabsf(p.x) <= kEps) & (unsigned)(fabsf(p.y) <= kEps) & (unsigned)(fabsf(p.z) <= kEps)) != 0u;
}

enum : uint32_t { LOC_VERTEX = 0, LOC_EDGE = 1, LOC_FACE = 2, LOC_SOLID = 3 };
struct Loc {
    uint32_t kind, id;
    float b0, b1, b2;
};

__device__ __forceinline__ f3 seg_origin(f3 a, f3 b, Loc &l) {
    const f3 ab = sub3(b, a);
    const f3 ap = neg3(a);  // origin - a
    const float ab_ap = dot3(ab, ap);
    const float sqnab = norm_sq3(ab);
    if (ab_ap <= 0.0f) {
        l = Loc{LOC_VERTEX, 0, 1.0f, 0.0f, 0.0f};
        return a;
    }
    if (ab_ap >= sqnab) {
        l = Loc{LOC_VERTEX, 1, 1.0f, 0.0f, 0.0f};
        return b;
    }
    const float u = ab_ap / sqnab;
    l = Loc{LOC_EDGE, 0, 1.0f - u, u, 0.0f};
    return add3(a, mul3(ab, u));
}

// Triangle projection of the origin (solid); *inside = relative_eq!(proj, origin).
__device__ __forceinline__ f3 tri_origin(f3 a, f3 b, f3 c, Loc &l, bool &inside) {
    const f3 ab = sub3(b, a), ac = sub3(c, a), ap = neg3(a);
    const float ab_ap = dot3(ab, ap), ac_ap = dot3(ac, ap);
    if (ab_ap <= 0.0f && ac_ap <= 0.0f) {
        l = Loc{LOC_VERTEX, 0, 0, 0, 0};
        inside = near_origin(a);
        return a;
    }
    const f3 bp = neg3(b);
    const float ab_bp = dot3(ab, bp), ac_bp = dot3(ac, bp);
    if (ab_bp >= 0.0f && ac_bp <= ab_bp) {
        l = Loc{LOC_VERTEX, 1, 0, 0, 0};
        inside = near_origin(b);
        return b;
    }
    const f3 cp = neg3(c);
    const float ab_cp = dot3(ab, cp), ac_cp = dot3(ac, cp);
    if (ac_cp >= 0.0f && ab_cp <= ac_cp) {
        l = Loc{LOC_VERTEX, 2, 0, 0, 0};
        inside = near_origin(c);
        return c;
    }
    const f3 bc = sub3(c, b);
    const f3 n = cross3(ab, ac);
    const float vc = dot3(n, cross3(ab, ap));
    if (vc < 0.0f && ab_ap >= 0.0f && ab_bp <= 0.0f) {
        const float v = ab_ap / norm_sq3(ab);
        l = Loc{LOC_EDGE, 0, 1.0f - v, v, 0.0f};
        const f3 r = add3(a, mul3(ab, v));
        inside = near_origin(r);
        return r;
    }
    const float vb = -dot3(n, cross3(ac, cp));
    if (vb < 0.0f && ac_ap >= 0.0f && ac_cp <= 0.0f) {
        const float w = ac_ap / norm_sq3(ac);
        l = Loc{LOC_EDGE, 2, 1.0f - w, w, 0.0f};
        const f3 r = add3(a, mul3(ac, w));
        inside = near_origin(r);
        return r;
    }
    const float va = dot3(n, cross3(bc, bp));
    if (va < 0.0f && ac_bp - ab_bp >= 0.0f && ab_cp - ac_cp >= 0.0f) {
        const float w = dot3(bc, bp) / norm_sq3(bc);
        l = Loc{LOC_EDGE, 1, 1.0f - w, w, 0.0f};
        const f3 r = add3(b, mul3(bc, w));
        inside = near_origin(r);
        return r;
    }
    const uint32_t side = dot3(n, ap) >= 0.0f ? 0u : 1u;
    if (va + vb + vc != 0.0f) {
        const float denom = 1.0f / (va + vb + vc);
        const float v = vb * denom, w = vc * denom;
        l = Loc{LOC_FACE, side, 1.0f - v - w, v, w};
        const f3 r = add3(add3(a, mul3(ab, v)), mul3(ac, w));
        inside = near_origin(r);
        return r;
    }
    l = Loc{LOC_SOLID, 0, 0, 0, 0};
    inside = true;
    return f3{0.0f, 0.0f, 0.0f};
}

// Tetrahedron projection of the origin (solid).  Not inlined (it is large and
// runs on the last GJK iterations only); inputs and outputs pass in registers.
struct TetOut {
    f3 p;
    Loc l;
};
__device__ __forceinline__ f3 tet_origin_impl(f3 a, f3 b, f3 c, f3 d, Loc &l);
__device__ __noinline__ TetOut tet_origin(f3 a, f3 b, f3 c, f3 d) {
    TetOut o;
    o.p = tet_origin_impl(a, b, c, d, o.l);
    return o;
}
__device__ __forceinline__ f3 tet_origin_impl(f3 a, f3 b, f3 c, f3 d, Loc &l) {
    const f3 ab = sub3(b, a), ac = sub3(c, a), ad = sub3(d, a), ap = neg3(a);
    const float ap_ab = dot3(ap, ab), ap_ac = dot3(ap, ac), ap_ad = dot3(ap, ad);
    if (ap_ab <= 0.0f && ap_ac <= 0.0f && ap_ad <= 0.0f) {
        l = Loc{LOC_VERTEX, 0, 0, 0, 0};
        return a;
    }
    const f3 bc = sub3(c, b), bd = sub3(d, b), bp = neg3(b);
    const float bp_bc = dot3(bp, bc), bp_bd = dot3(bp, bd), bp_ab = dot3(bp, ab);
    if (bp_bc <= 0.0f && bp_bd <= 0.0f && bp_ab >= 0.0f) {
        l = Loc{LOC_VERTEX, 1, 0, 0, 0};
        return b;
    }
    const f3 cd = sub3(d, c), cp = neg3(c);
    const float cp_ac = dot3(cp, ac), cp_bc = dot3(cp, bc), cp_cd = dot3(cp, cd);
    if (cp_cd <= 0.0f && cp_bc >= 0.0f && cp_ac >= 0.0f) {
        l = Loc{LOC_VERTEX, 2, 0, 0, 0};
        return c;
    }
    const f3 dp = neg3(d);
    const float dp_cd = dot3(dp, cd), dp_bd = dot3(dp, bd), dp_ad = dot3(dp, ad);
    if (dp_ad >= 0.0f && dp_bd >= 0.0f && dp_cd >= 0.0f) {
        l = Loc{LOC_VERTEX, 3, 0, 0, 0};
        return d;
    }
    // Edge regions; each test also yields the two signed volumes the face
    // tests reuse.
    f3 r;
#define PQ_EDGE(I, EA, NABC, NABD, EAP, EAB, EAP_AB, EBP_AB, D1, D2)                                 \
    float D1, D2;                                                                                    \
    {                                                                                                \
        const float ab_ab = (EAP_AB) - (EBP_AB);                                                     \
        const f3 x = cross3(EAP, EAB);                                                               \
        D1 = dot3(x, NABC);                                                                          \
        D2 = dot3(x, NABD);                                                                          \
        if (ab_ab != 0.0f && D1 >= 0.0f && D2 >= 0.0f && (EAP_AB) >= 0.0f && (EAP_AB) <= ab_ab) {    \
            const float u = (EAP_AB) / ab_ab;                                                        \
            l = Loc{LOC_EDGE, I, 1.0f - u, u, 0.0f};                                                 \
            return add3(EA, mul3(EAB, u));                                                           \
        }                                                                                            \
    }
    const f3 nabc = cross3(ab, ac), nabd = cross3(ab, ad);
    PQ_EDGE(0, a, nabc, nabd, ap, ab, ap_ab, bp_ab, dabc, dabd)
    const f3 nacd = cross3(ac, ad);
    PQ_EDGE(1, a, nacd, neg3(nabc), ap, ac, ap_ac, cp_ac, dacd, dacb)
    PQ_EDGE(2, a, neg3(nabd), neg3(nacd), ap, ad, ap_ad, dp_ad, dadb, dadc)
    const f3 nbcd = cross3(bc, bd);
    PQ_EDGE(3, b, nabc, nbcd, bp, bc, bp_bc, cp_bc, dbca, dbcd)
    PQ_EDGE(4, b, neg3(nbcd), nabd, bp, bd, bp_bd, dp_bd, dbdc, dbda)
    PQ_EDGE(5, c, nacd, nbcd, cp, cd, cp_cd, dp_cd, dcda, dcdb)
#undef PQ_EDGE
#define PQ_FACE(I, FA, FB, FC, FAP, FBP, FCP, FAB, FAC, FAD, D1, D2, D3)                            \
    if (D1 < 0.0f && D2 < 0.0f && D3 < 0.0f) {                                                      \
        const f3 n = cross3(FAB, FAC);                                                              \
        if (dot3(n, FAD) * dot3(n, FAP) < 0.0f) {                                                   \
            const float nn = sqrtf(norm_sq3(n));                                                    \
            if (!(nn <= kEps)) {                                                                    \
                const f3 nu = div3(n, nn);                                                          \
                const float vc = dot3(nu, cross3(FAP, FBP));                                        \
                const float va = dot3(nu, cross3(FBP, FCP));                                        \
                const float vb = dot3(nu, cross3(FCP, FAP));                                        \
                const float inv = 1.0f / ((va + vb) + vc);                                          \
                const float b0 = va * inv, b1 = vb * inv, b2 = vc * inv;                            \
                l = Loc{LOC_FACE, I, b0, b1, b2};                                                   \
                return add3(add3(mul3(FA, b0), mul3(FB, b1)), mul3(FC, b2));                        \
            }                                                                                       \
        }                                                                                           \
    }
    PQ_FACE(0, a, b, c, ap, bp, cp, ab, ac, ad, dabc, dbca, dacb)
    PQ_FACE(1, a, b, d, ap, bp, dp, ab, ad, ac, dadb, dabd, dbda)
    PQ_FACE(2, a, c, d, ap, cp, dp, ac, ad, ab, dacd, dcda, dadc)
    PQ_FACE(3, b, c, d, bp, cp, dp, bc, bd, neg3(ab), dbcd, dcdb, dbdc)
#undef PQ_FACE
    (void)r;
    l = Loc{LOC_SOLID, 0, 0, 0, 0};
    return f3{0.0f, 0.0f, 0.0f};
}

// VoronoiSimplex: vertices, barycentric coordinates of the last projection and
// the previous simplex (as a permutation of the current vertex slots).
struct Simplex {
    Cso v[4];
    float proj[3];
    float prev_proj[3];
    uint32_t perm;  // prev_vertices, 4 x 8 bits
    uint32_t dim, prev_dim;
};

__device__ __forceinline__ uint32_t perm_get(uint32_t perm, uint32_t i) { return (perm >> (8 * i)) & 0xffu; }
__device__ __forceinline__ uint32_t perm_swap(uint32_t perm, uint32_t i, uint32_t j) {
    const uint32_t a = perm_get(perm, i), b = perm_get(perm, j);
    perm &= ~((0xffu << (8 * i)) | (0xffu << (8 * j)));
    return perm | (b << (8 * i)) | (a << (8 * j));
}
constexpr uint32_t kPermIdentity = 0x03020100u;

// Slot access by a run-time index through selects, so the simplex stays in
// registers (an indexed private array would live in scratch memory).
// (Selecting whole structs with ?: would select an address: component-wise
// scalar selects keep everything in registers.)
__device__ __forceinline__ float sel4(float a, float b, float c, float d, uint32_t i) {
    return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}
__device__ __forceinline__ f3 sel4(const f3 &a, const f3 &b, const f3 &c, const f3 &d, uint32_t i) {
    return f3{sel4(a.x, b.x, c.x, d.x, i), sel4(a.y, b.y, c.y, d.y, i), sel4(a.z, b.z, c.z, d.z, i)};
}
__device__ __forceinline__ Cso sx_get(const Simplex &s, uint32_t i) {
    return Cso{sel4(s.v[0].p, s.v[1].p, s.v[2].p, s.v[3].p, i), sel4(s.v[0].o1, s.v[1].o1, s.v[2].o1, s.v[3].o1, i),
               sel4(s.v[0].o2, s.v[1].o2, s.v[2].o2, s.v[3].o2, i)};
}
__device__ __forceinline__ f3 sel2(bool c, const f3 &a, const f3 &b) {
    return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}
__device__ __forceinline__ Cso csel(bool c, const Cso &a, const Cso &b) {
    return Cso{sel2(c, a.p, b.p), sel2(c, a.o1, b.o1), sel2(c, a.o2, b.o2)};
}
// Every slot is rewritten (with a value select), so no store goes through a
// run-time-selected address.
__device__ __forceinline__ void sx_set(Simplex &s, uint32_t i, const Cso &c) {
    s.v[0] = csel(i == 0, c, s.v[0]);
    s.v[1] = csel(i == 1, c, s.v[1]);
    s.v[2] = csel(i == 2, c, s.v[2]);
    s.v[3] = csel(i == 3, c, s.v[3]);
}

__device__ __forceinline__ void sx_swap(Simplex &s, uint32_t i, uint32_t j) {
    const Cso a = sx_get(s, i), b = sx_get(s, j);
    sx_set(s, i, b);
    sx_set(s, j, a);
    s.perm = perm_swap(s.perm, i, j);
}

__device__ __forceinline__ void sx_reset(Simplex &s, const Cso &pt) {
    s.dim = 0;
    s.prev_dim = 0;
    s.v[0] = pt;
    s.perm = kPermIdentity;
    s.proj[0] = s.proj[1] = s.proj[2] = 0.0f;
    s.prev_proj[0] = s.prev_proj[1] = s.prev_proj[2] = 0.0f;
}

// VoronoiSimplex::add_point (voronoi_simplex3.rs:57-92).
__device__ __forceinline__ bool sx_add(Simplex &s, const Cso &pt) {
    s.prev_dim = s.dim;
    s.prev_proj[0] = s.proj[0];
    s.prev_proj[1] = s.proj[1];
    s.prev_proj[2] = s.proj[2];
    s.perm = kPermIdentity;
    if (s.dim == 0) {
        if (norm_sq3(sub3(s.v[0].p, pt.p)) < kEpsTol) return false;
    } else if (s.dim == 1) {
        const f3 ab = sub3(s.v[1].p, s.v[0].p), ac = sub3(pt.p, s.v[0].p);
        if (norm_sq3(cross3(ab, ac)) < kEpsTol) return false;
    } else {
        const f3 ab = sub3(s.v[1].p, s.v[0].p), ac = sub3(s.v[2].p, s.v[0].p), ap = sub3(pt.p, s.v[0].p);
        const f3 c = cross3(ab, ac);
        const f3 n = div3(c, sqrtf(norm_sq3(c)));
        if (fabsf(dot3(n, ap)) < kEpsTol) return false;
    }
    s.dim += 1;
    sx_set(s, s.dim, pt);
    return true;
}

// VoronoiSimplex::project_origin_and_reduce (voronoi_simplex3.rs:123-279).
__device__ __forceinline__ f3 sx_project(Simplex &s) {
    if (s.dim == 0) {
        s.proj[0] = 1.0f;
        return s.v[0].p;
    }
    Loc l;
    if (s.dim == 1) {
        const f3 p = seg_origin(s.v[0].p, s.v[1].p, l);
        if (l.kind == LOC_VERTEX) {
            if (l.id == 1) sx_swap(s, 0, 1);
            s.proj[0] = 1.0f;
            s.dim = 0;
        } else {
            s.proj[0] = l.b0;
            s.proj[1] = l.b1;
        }
        return p;
    }
    if (s.dim == 2) {
        bool inside;
        const f3 p = tri_origin(s.v[0].p, s.v[1].p, s.v[2].p, l, inside);
        if (l.kind == LOC_VERTEX) {
            sx_swap(s, 0, l.id);
            s.proj[0] = 1.0f;
            s.dim = 0;
        } else if (l.kind == LOC_EDGE) {
            if (l.id == 0) {
                s.proj[0] = l.b0;
                s.proj[1] = l.b1;
            } else if (l.id == 1) {
                sx_swap(s, 0, 2);
                s.proj[0] = l.b1;
                s.proj[1] = l.b0;
            } else {
                sx_swap(s, 1, 2);
                s.proj[0] = l.b0;
                s.proj[1] = l.b1;
            }
            s.dim = 1;
        } else if (l.kind == LOC_FACE) {
            s.proj[0] = l.b0;
            s.proj[1] = l.b1;
            s.proj[2] = l.b2;
        }
        return p;
    }
    const TetOut to = tet_origin(s.v[0].p, s.v[1].p, s.v[2].p, s.v[3].p);
    const f3 p = to.p;
    l = to.l;
    if (l.kind == LOC_VERTEX) {
        sx_swap(s, 0, l.id);
        s.proj[0] = 1.0f;
        s.dim = 0;
    } else if (l.kind == LOC_EDGE) {
        switch (l.id) {
            case 1: sx_swap(s, 1, 2); break;
            case 2: sx_swap(s, 1, 3); break;
            case 3: sx_swap(s, 0, 2); break;
            case 4: sx_swap(s, 0, 3); break;
            case 5: sx_swap(s, 0, 2); sx_swap(s, 1, 3); break;
            default: break;
        }
        const bool flip = l.id == 3 || l.id == 4;
        s.proj[0] = flip ? l.b1 : l.b0;
        s.proj[1] = flip ? l.b0 : l.b1;
        s.dim = 1;
    } else if (l.kind == LOC_FACE) {
        // face 0 keeps (v0, v1, v2); faces 1, 2, 3 move v3 into slot 2, 1, 0
        if (l.id != 0) sx_set(s, 3 - l.id, s.v[3]);
        if (l.id <= 1) {
            s.proj[0] = l.b0; s.proj[1] = l.b1; s.proj[2] = l.b2;
        } else if (l.id == 2) {
            s.proj[0] = l.b0; s.proj[1] = l.b2; s.proj[2] = l.b1;
        } else {
            s.proj[0] = l.b2; s.proj[1] = l.b0; s.proj[2] = l.b1;
        }
        s.dim = 2;
    }
    return p;
}

// gjk::result (gjk.rs:797-818): points from the current or previous simplex.
__device__ __forceinline__ void sx_points(const Simplex &s, bool prev, f3 &p1, f3 &p2) {
    p1 = p2 = f3{0.0f, 0.0f, 0.0f};
    const uint32_t k = prev ? s.prev_dim : s.dim;  // at most 2 here
#pragma unroll
    for (uint32_t i = 0; i < 3; ++i) {
        if (i > k) break;
        const float c = prev ? s.prev_proj[i] : s.proj[i];
        const Cso pt = sx_get(s, prev ? perm_get(s.perm, i) : i);
        p1 = add3(p1, mul3(pt.o1, c));
        p2 = add3(p2, mul3(pt.o2, c));
    }
}

enum : uint32_t { GJK_INTERSECTION = 0, GJK_CLOSEST = 1, GJK_NO_INTERSECTION = 2 };

struct GjkOut {
    uint32_t kind;
    uint32_t iters;
    f3 p1, p2, n;
};

// gjk::closest_points (gjk.rs:353-453) with exact_dist = true.
__device__ __forceinline__ GjkOut gjk_closest_points(const iso &pos12, const ShapeRef &g1, const ShapeRef &g2,
                                                  float max_dist, Simplex &s) {
    const float eps_rel = sqrtf(kEpsTol);
    GjkOut out{GJK_INTERSECTION, 0, f3{0, 0, 0}, f3{0, 0, 0}, f3{0, 0, 0}};
    f3 proj = sx_project(s);
    f3 old_dir;
    {
        const float sq = norm_sq3(proj);
        if (!(sq > 0.0f)) return out;
        old_dir = neg3(div3(proj, sqrtf(sq)));
    }
    float max_bound = kFltMax;
    // One exit: each outcome records its kind, normal and which simplex gives
    // the points, then leaves the loop; the points are taken once below.  With
    // a return per outcome the compiler nested the loop once per exit path.
    enum : uint32_t { kNoPoints = 0, kCurrent = 1, kPrevious = 2 };
    uint32_t points = kNoPoints;
    for (uint32_t niter = 0;;) {
        const float old_max_bound = max_bound;
        const f3 mp = neg3(proj);
        const float sq = norm_sq3(mp);
        if (!(sq > kEpsTol * kEpsTol)) {
            out.kind = GJK_INTERSECTION;
            break;
        }
        const float nrm = sqrtf(sq);
        const f3 dir = div3(mp, nrm);
        max_bound = nrm;
        if (max_bound >= old_max_bound) {
            out.kind = GJK_CLOSEST;
            out.n = old_dir;
            points = kPrevious;
            break;
        }
        const Cso c = cso_from_shapes(pos12, g1, g2, dir);
        const float min_bound = -dot3(dir, c.p);
        if (!isfinite(min_bound)) {  // parry asserts here (gjk.rs:407)
            out.kind = GJK_NO_INTERSECTION;
            out.n = f3{1.0f, 0.0f, 0.0f};
            break;
        }
        if (min_bound > max_dist) {
            out.kind = GJK_NO_INTERSECTION;
            out.n = dir;
            break;
        }
        if (max_bound - min_bound <= eps_rel * max_bound || !sx_add(s, c)) {
            out.kind = GJK_CLOSEST;
            out.n = dir;
            points = kCurrent;
            break;
        }
        old_dir = dir;
        proj = sx_project(s);
        if (s.dim == 3) {
            if (min_bound >= kEpsTol) {
                out.kind = GJK_CLOSEST;
                out.n = old_dir;
                points = kPrevious;
            } else {
                out.kind = GJK_INTERSECTION;
            }
            break;
        }
        ++niter;
        out.iters = niter;
        if (niter == kGjkMaxIters) {
            out.kind = GJK_NO_INTERSECTION;
            out.n = f3{1.0f, 0.0f, 0.0f};
            break;
        }
    }
    if (points != kNoPoints) sx_points(s, points == kPrevious, out.p1, out.p2);
    return out;
}


// ---------------------------------------------------------------------------
// Ball against a convex shape (contact_ball_convex_polyhedron.rs:12-63): the
// point projections (solid = false) of the ball centre and the contact built
// from them.  A ConvexPolyhedron is projected by GJK (and EPA when the centre
// is inside) against ConstantOrigin; the caller runs those.
// ---------------------------------------------------------------------------
struct PointProj {
    bool inside;
    f3 point;
};
enum : uint32_t { CF_UNKNOWN = 0, CF_VERTEX = 1, CF_EDGE = 2, CF_FACE = 3 };

// approx relative_eq!(a, b) with epsilon = max_relative = f32::EPSILON.
__device__ __forceinline__ bool rel_eq1(float a, float b) {
    if (a == b) return true;
    if (isinf(a) || isinf(b)) return false;
    const float d = fabsf(a - b);
    if (d <= kEps) return true;
    const float fa = fabsf(a), fb = fabsf(b);
    const float largest = fb > fa ? fb : fa;
    return d <= largest * kEps;
}
__device__ __forceinline__ bool rel_eq3(f3 a, f3 b) { return rel_eq1(a.x, b.x) && rel_eq1(a.y, b.y) && rel_eq1(a.z, b.z); }

// Unit::try_new_and_get(v, EPS): |v|^2 > EPS^2.
__device__ __forceinline__ bool try_unit(f3 v, f3 &u, float &len) {
    const float sq = norm_sq3(v);
    if (!(sq > kEps * kEps)) return false;
    len = sqrtf(sq);
    u = div3(v, len);
    return true;
}

__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ void set_comp(f3 &v, int i, float x) {
    if (i == 0) v.x = x;
    else if (i == 1) v.y = x;
    else v.z = x;
}

// Segment::project_local_point (point_segment.rs:45-85) with relative_eq!(proj, pt).
__device__ __forceinline__ PointProj segment_point(f3 a, f3 b, f3 pt) {
    const f3 ab = sub3(b, a), ap = sub3(pt, a);
    const float ab_ap = dot3(ab, ap), sqnab = norm_sq3(ab);
    f3 p;
    if (ab_ap <= 0.0f) p = a;
    else if (ab_ap >= sqnab) p = b;
    else p = add3(a, mul3(ab, ab_ap / sqnab));
    return PointProj{rel_eq3(p, pt), p};
}

// Aabb(-he, he)::project_local_point_and_get_feature (point_aabb.rs:9-132).
__device__ PointProj cuboid_point_feature(const ShapeRef &s, f3 pt, uint32_t &fk, uint32_t &fid) {
    const f3 maxs = f3{s.p0, s.p1, s.p2}, mins = neg3(maxs);
    const f3 mins_pt = sub3(mins, pt), pt_maxs = sub3(pt, maxs);
    const f3 zero = f3{0.0f, 0.0f, 0.0f};
    f3 shift = sub3(max3(mins_pt, zero), max3(pt_maxs, zero));
    const bool inside = shift.x == 0.0f && shift.y == 0.0f && shift.z == 0.0f;
    if (inside) {
        float best = -kFltMax;
        bool is_mins = false;
        int best_id = 0;
        for (int i = 0; i < 3; ++i) {
            const float a = comp(mins_pt, i), b = comp(pt_maxs, i);
            if (a < b) {
                if (b > best) {
                    best_id = i;
                    is_mins = false;
                    best = b;
                }
            } else if (a > best) {
                best_id = i;
                is_mins = true;
                best = a;
            }
        }
        shift = zero;
        set_comp(shift, best_id, is_mins ? best : -best);
    }
    const f3 ls = add3(pt, shift);
    int nzero = 0, last_zero = 0, last_nz = 0;
    for (int i = 0; i < 3; ++i) {
        if (comp(shift, i) == 0.0f) {
            ++nzero;
            last_zero = i;
        } else {
            last_nz = i;
        }
    }
    const f3 center = mul3(add3(mins, maxs), 0.5f);
    fk = CF_UNKNOWN;
    fid = 0;
    if (nzero == 3) {
        for (int i = 0; i < 3; ++i) {
            if (comp(ls, i) > comp(maxs, i) - kEps) {
                fk = CF_FACE;
                fid = (uint32_t)i;
                break;
            }
            if (comp(ls, i) <= comp(mins, i) + kEps) {
                fk = CF_FACE;
                fid = (uint32_t)(i + 3);
                break;
            }
        }
    } else if (nzero == 2) {
        fk = CF_FACE;
        fid = comp(ls, last_nz) < comp(center, last_nz) ? (uint32_t)(last_nz + 3) : (uint32_t)last_nz;
    } else {
        uint32_t id = 0;
        for (int i = 0; i < 3; ++i)
            if (comp(ls, i) < comp(center, i)) id |= 1u << i;
        if (nzero == 0) {
            fk = CF_VERTEX;
            fid = id;
        } else {
            fk = CF_EDGE;
            fid = (id << 2) | (uint32_t)last_zero;
        }
    }
    return PointProj{inside, ls};
}

// Cuboid::feature_normal (cuboid.rs:401-448).
__device__ bool cuboid_feature_normal(uint32_t fk, uint32_t id, f3 &n) {
    n = f3{0.0f, 0.0f, 0.0f};
    if (fk == CF_FACE) {
        if (id < 3) set_comp(n, (int)id, 1.0f);
        else set_comp(n, (int)id - 3, -1.0f);
        return true;
    }
    if (fk == CF_EDGE) {
        const uint32_t edge = id & 3u, f1 = (edge + 1) % 3, f2 = (edge + 2) % 3, signs = id >> 2;
        set_comp(n, (int)f1, (signs & (1u << f1)) ? -1.0f : 1.0f);
        set_comp(n, (int)f2, (signs & (1u << f2)) ? -1.0f : 1.0f);
        n = div3(n, sqrtf(norm_sq3(n)));
        return true;
    }
    if (fk == CF_VERTEX) {
        for (int i = 0; i < 3; ++i) set_comp(n, i, (id & (1u << i)) ? -1.0f : 1.0f);
        n = div3(n, sqrtf(norm_sq3(n)));
        return true;
    }
    return false;
}

// Capsule::project_local_point (point_capsule.rs:9-50), solid = false.
__device__ PointProj capsule_point(const ShapeRef &s, f3 pt) {
    const f3 a = f3{0.0f, -s.p0, 0.0f}, b = f3{0.0f, s.p0, 0.0f};
    const float r = s.p1;
    const PointProj proj = segment_point(a, b, pt);
    f3 dir;
    float dist;
    if (try_unit(sub3(pt, proj.point), dir, dist)) return PointProj{dist <= r, add3(proj.point, mul3(dir, r))};
    f3 sd;
    float sl;
    if (try_unit(sub3(b, a), sd, sl)) {
        // WBasis::orthonormal_basis()[0] (utils/wops.rs:96-110)
        const float sign = copysignf(1.0f, sd.z);
        const float aa = -1.0f / (sign + sd.z);
        const float bb = sd.x * sd.y * aa;
        const f3 e0 = f3{1.0f + sign * sd.x * sd.x * aa, sign * bb, -sign * sd.x};
        return PointProj{true, add3(proj.point, mul3(e0, r))};
    }
    return PointProj{true, add3(proj.point, f3{0.0f, r, 0.0f})};
}

// Cone::project_local_point (point_cone.rs:8-63), solid = false.
__device__ PointProj cone_point(const ShapeRef &s, f3 pt) {
    const float h = s.p0, r = s.p1;
    float dx = pt.x, dz = pt.z;
    const float planar = sqrtf(dx * dx + dz * dz);
    dx = dx / planar;
    dz = dz / planar;
    if (planar <= kEps) {
        dx = 1.0f;
        dz = 0.0f;
    }
    const f3 on_basis = f3{pt.x, -h, pt.z};
    if (pt.y < -h && planar <= r) return PointProj{false, on_basis};
    const f3 circle = f3{dx * r, -h, dz * r};
    const f3 apex = f3{0.0f, h, 0.0f};
    const f3 side_dir = sub3(circle, apex);
    PointProj proj = segment_point(apex, circle, pt);
    const f3 apex_to_basis = f3{0.0f, -2.0f * h, 0.0f};
    if (pt.y >= -h && pt.y <= h &&
        dot3(cross3(side_dir, sub3(pt, apex)), cross3(side_dir, apex_to_basis)) >= 0.0f) {
        const f3 d1 = sub3(proj.point, pt), d2 = sub3(on_basis, pt);
        if (norm_sq3(d1) > norm_sq3(d2)) return PointProj{true, on_basis};
        proj.inside = true;
    }
    return proj;
}

// The isometries of local_point_projection_on_support_map (point_support_map.rs:26-29):
// the initial CSO point uses m_inv = (I, c), GJK/EPA run on m.inverse() with m = (I, -c).
__device__ __forceinline__ iso point_proj_seed_iso(f3 c) { return iso{quat{0.0f, 0.0f, 0.0f, 1.0f}, c}; }
__device__ __forceinline__ iso point_proj_iso(f3 c) {
    return iso_inverse(iso{quat{0.0f, 0.0f, 0.0f, 1.0f}, neg3(c)});
}
__device__ __forceinline__ ShapeRef origin_shape() { return ShapeRef{kShapeOrigin, 0, nullptr, 0.0f, 0.0f, 0.0f}; }

// Triangle::normal (triangle.rs:226-228): the unit scaled normal, if not degenerate.
__device__ __forceinline__ bool triangle_normal(const ShapeRef &t, f3 &n) {
    const float4 a = t.v[0], b = t.v[1], c = t.v[2];
    const f3 pa = f3{a.x, a.y, a.z};
    float len;
    return try_unit(cross3(sub3(f3{b.x, b.y, b.z}, pa), sub3(f3{c.x, c.y, c.z}, pa)), n, len);
}

// Triangle::project_local_point_and_get_location (point_triangle.rs:54-291), 3D,
// for any point.  `inside` is PointProjection::is_inside: relative_eq!
// (projection, point) in the Voronoi-region branches, true in a degenerate
// face's fallbacks (the point itself when solid, else the closest edge).
__device__ __forceinline__ f3 tri_project(f3 a, f3 b, f3 c, f3 pt, bool solid, bool &inside) {
    const f3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(pt, a);
    const float ab_ap = dot3(ab, ap), ac_ap = dot3(ac, ap);
    if (ab_ap <= 0.0f && ac_ap <= 0.0f) {
        inside = rel_eq3(a, pt);
        return a;
    }
    const f3 bp = sub3(pt, b);
    const float ab_bp = dot3(ab, bp), ac_bp = dot3(ac, bp);
    if (ab_bp >= 0.0f && ac_bp <= ab_bp) {
        inside = rel_eq3(b, pt);
        return b;
    }
    const f3 cp = sub3(pt, c);
    const float ab_cp = dot3(ab, cp), ac_cp = dot3(ac, cp);
    if (ac_cp >= 0.0f && ab_cp <= ac_cp) {
        inside = rel_eq3(c, pt);
        return c;
    }
    const f3 bc = sub3(c, b);
    // stable_check_edges_voronoi (dim3)
    const f3 n = cross3(ab, ac);
    const float vc = dot3(n, cross3(ab, ap));
    f3 r;
    if (vc < 0.0f && ab_ap >= 0.0f && ab_bp <= 0.0f) {
        r = add3(a, mul3(ab, ab_ap / norm_sq3(ab)));
        inside = rel_eq3(r, pt);
        return r;
    }
    const float vb = -dot3(n, cross3(ac, cp));
    if (vb < 0.0f && ac_ap >= 0.0f && ac_cp <= 0.0f) {
        r = add3(a, mul3(ac, ac_ap / norm_sq3(ac)));
        inside = rel_eq3(r, pt);
        return r;
    }
    const float va = dot3(n, cross3(bc, bp));
    if (va < 0.0f && ac_bp - ab_bp >= 0.0f && ab_cp - ac_cp >= 0.0f) {
        r = add3(b, mul3(bc, dot3(bc, bp) / norm_sq3(bc)));
        inside = rel_eq3(r, pt);
        return r;
    }
    if (va + vb + vc != 0.0f) {
        const float denom = 1.0f / (va + vb + vc);
        const float v = vb * denom, w = vc * denom;
        r = add3(add3(a, mul3(ab, v)), mul3(ac, w));
        inside = rel_eq3(r, pt);
        return r;
    }
    inside = true;
    if (solid) return pt;
    const float v = ab_ap / (ab_ap - ab_bp);
    const float w = ac_ap / (ac_ap - ac_cp);
    const float u = (ac_bp - ab_bp) / (ac_bp - ab_bp + ab_cp - ac_cp);
    const float d_ab = norm_sq3(ap) - (norm_sq3(ab) * v * v);
    const float d_ac = norm_sq3(ap) - (norm_sq3(ac) * w * w);
    const float d_bc = norm_sq3(bp) - (norm_sq3(bc) * u * u);
    if (d_ab < d_ac) return d_ab < d_bc ? add3(a, mul3(ab, v)) : add3(b, mul3(bc, u));
    return d_ac < d_bc ? add3(a, mul3(ac, w)) : add3(b, mul3(bc, u));
}

// Triangle::project_local_point_and_get_feature: solid = true in 3D (point_triangle.rs:26-48).
__device__ __forceinline__ PointProj triangle_point(const ShapeRef &t, f3 pt) {
    const float4 a = t.v[0], b = t.v[1], c = t.v[2];
    bool inside = false;
    const f3 p = tri_project(f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, pt, true, inside);
    return PointProj{inside, p};
}

// contact_convex_polyhedron_ball (contact_ball_convex_polyhedron.rs:27-63) in the
// convex shape's frame: pos12 maps the ball's frame into the convex shape's.
// `has_fn`/`fnrm`: the shape's feature_normal_at_point for the projection's
// feature, when it has one (a Cuboid's face/edge/vertex normal, a Triangle's
// normal); used only when the ball's centre lies on the projection.
// Returns false for None.
__device__ bool convex_ball_contact(const iso &pos12, float radius, float prediction, const PointProj &proj,
                                    bool has_fn, f3 fnrm, f3 &p1, f3 &p2, f3 &n1, f3 &n2, float &dist) {
    const f3 c = pos12.t;
    f3 dir1;
    float len;
    if (try_unit(sub3(proj.point, c), dir1, len)) {
        if (proj.inside) {
            dist = -len - radius;
            n1 = dir1;
        } else {
            dist = len - radius;
            n1 = neg3(dir1);
        }
    } else {
        // A ConvexPolyhedron's feature is Unknown whenever this branch is taken
        // (same threshold, point_support_map.rs:66-72); Capsule and Cone have no
        // feature normal (shape.rs:425-431).
        dist = -radius;
        if (has_fn) {
            n1 = fnrm;
        } else {
            f3 u;
            float l;
            n1 = try_unit(proj.point, u, l) ? u : f3{0.0f, 1.0f, 0.0f};
        }
    }
    if (!(dist <= prediction)) return false;
    n2 = qrot_inv(pos12.q, neg3(n1));
    p2 = mul3(n2, radius);
    p1 = proj.point;
    return true;
}

}  // namespace pq
