// manifold.hip -- contact manifolds between polygonal feature maps (gfx950).
//
// contact_manifold_pfm_pfm (src/query/contact_manifolds/contact_manifolds_pfm_pfm.rs:44-162)
// for a fresh manifold (no points and a zero cached normal: try_update_contacts_eps
// fails at once, contact_manifold.rs:662-670, GJK takes parry's default seed, and
// match_contacts has nothing to match) and no normal constraints:
//   1. the batched contact kernels in their feature-map mode: GJK/EPA on the two
//      feature maps (a Capsule is its segment) with prediction + both border
//      radii, results in local frames;
//   2. one lane per pair: the support features of both maps along the contact
//      normal (polygonal_feature_map.rs:30-154, cuboid.rs:267-352,
//      triangle.rs:280-282, convex_polyhedron.rs:959-992 from the faces given
//      by pq_shapes_set_polyhedron_faces), their clipping (PolygonalFeature::contacts,
//      polygonal_feature3d.rs:74-381: edge-edge or face-face in the plane of the
//      normal's orthonormal basis, utils/wops.rs:96-109), GJK's own point last,
//      and the border radii applied.
#include "contact_gjk.h"
#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBlock = 256;
constexpr float kCosFracPi8 = 0.92387953251f;  // utils/consts.rs:11
constexpr float kCos1Degrees = 0.99984769515f;  // utils/consts.rs:9
constexpr uint32_t kPfVertex = 1u << 30, kPfEdge = 2u << 30, kPfFace = 3u << 30;  // feature_id.rs:466-471

struct Feature {
    f3 v[4];
    uint32_t vids[4], eids[4], fid;
    int n;
};

struct P2 {
    float x, y;
};
__device__ __forceinline__ P2 sub2(P2 a, P2 b) { return P2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float dot2(P2 a, P2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float perp2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }

// approx::ulps_eq!(a, b) with epsilon = f32::EPSILON and max_ulps = 4.
__device__ __forceinline__ bool ulps_eq(float a, float b) {
    const float d = a > b ? a - b : b - a;
    if (d <= kEps) return true;
    if (signbit(a) != signbit(b) || isnan(a) || isnan(b)) return false;
    const uint32_t ia = __float_as_uint(a), ib = __float_as_uint(b);
    return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}

__device__ __forceinline__ float comp3(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

__device__ void segment_feature(float h, Feature &f) {
    const f3 a{0.0f, -h, 0.0f}, b{0.0f, h, 0.0f};
    f.v[0] = a;
    f.v[1] = b;
    f.v[2] = b;
    f.v[3] = b;
    for (int k = 0; k < 4; ++k) {
        f.vids[k] = kPfVertex | (k == 0 ? 0u : 1u);
        f.eids[k] = kPfEdge | 0u;
    }
    f.fid = kPfFace | 0u;
    f.n = 2;
}

__device__ void triangle_feature(const ShapeRef &t, Feature &f) {
    for (int k = 0; k < 3; ++k) {
        const float4 p = t.v[k];
        f.v[k] = f3{p.x, p.y, p.z};
    }
    f.v[3] = f.v[2];
    for (int k = 0; k < 4; ++k) {
        const uint32_t id = k < 3 ? (uint32_t)k : 2u;
        f.vids[k] = kPfVertex | id;
        f.eids[k] = kPfEdge | id;
    }
    f.fid = kPfFace | 0u;
    f.n = 3;
}

// Cuboid::support_face (cuboid.rs:267-352).
__device__ void cuboid_feature(f3 he, f3 d, Feature &f) {
    int ia = 0;
    float m = fabsf(d.x);
    if (fabsf(d.y) > m) {
        m = fabsf(d.y);
        ia = 1;
    }
    if (fabsf(d.z) > m) ia = 2;
    const float s = copysignf(1.0f, comp3(d, ia));
    if (ia == 0) {
        f.v[0] = f3{he.x * s, he.y, he.z};
        f.v[1] = f3{he.x * s, -he.y, he.z};
        f.v[2] = f3{he.x * s, -he.y, -he.z};
        f.v[3] = f3{he.x * s, he.y, -he.z};
    } else if (ia == 1) {
        f.v[0] = f3{he.x, he.y * s, he.z};
        f.v[1] = f3{-he.x, he.y * s, he.z};
        f.v[2] = f3{-he.x, he.y * s, -he.z};
        f.v[3] = f3{he.x, he.y * s, -he.z};
    } else {
        f.v[0] = f3{he.x, he.y, he.z * s};
        f.v[1] = f3{he.x, -he.y, he.z * s};
        f.v[2] = f3{-he.x, -he.y, he.z * s};
        f.v[3] = f3{-he.x, he.y, he.z * s};
    }
    const int si = s > 0.0f ? 1 : 0;
    // Vertex ids (2 x the sign bits of the vertex) and edge ids of each face, by
    // axis and sign index (cuboid.rs:297-337).
    constexpr uint32_t kVids[3][2][4] = {{{0, 4, 6, 2}, {8, 12, 14, 10}},
                                         {{0, 8, 10, 2}, {4, 12, 14, 6}},
                                         {{0, 4, 12, 8}, {2, 6, 14, 10}}};
    constexpr uint32_t kEids[3][2][4] = {
        {{0b11010000, 0b11011010, 0b11011001, 0b11001000}, {0b11110100, 0b11111110, 0b11111101, 0b11101100}},
        {{0b11100000, 0b11101100, 0b11101001, 0b11001000}, {0b11110010, 0b11111110, 0b11111011, 0b11011010}},
        {{0b11010000, 0b11110010, 0b11110100, 0b11100000}, {0b11011001, 0b11111011, 0b11111101, 0b11101001}}};
    for (int k = 0; k < 4; ++k) {
        f.vids[k] = kPfVertex | kVids[ia][si][k];
        f.eids[k] = kPfEdge | kEids[ia][si][k];
    }
    f.fid = kPfFace | (uint32_t)(ia + si * 3 + 10);
    f.n = 4;
}

// Cone's feature map (polygonal_feature_map.rs:108-154).
__device__ void cone_feature(float h, float r, f3 d, Feature &f) {
    float dx = 1.0f, dy = 0.0f;
    {
        const float nrm = sqrtf(d.x * d.x + d.z * d.z);  // Vector2::try_normalize(EPS)
        if (nrm > kEps) {
            dx = d.x / nrm;
            dy = d.z / nrm;
        }
    }
    for (int k = 0; k < 4; ++k) f.v[k] = f3{0.0f, 0.0f, 0.0f};
    if (d.y > 0.0f) {
        f.v[0] = f3{dx * r, -h, dy * r};
        f.v[1] = f3{0.0f, h, 0.0f};
        for (int k = 0; k < 4; ++k) {
            f.eids[k] = kPfEdge | 0u;
            f.vids[k] = kPfVertex | (k == 0 ? 1u : 11u);
        }
        f.fid = kPfFace | 0u;
        f.n = 2;
    } else {
        const float y = -h;
        f.v[0] = f3{dx * r, y, dy * r};
        f.v[1] = f3{-dy * r, y, dx * r};
        f.v[2] = f3{-dx * r, y, -dy * r};
        f.v[3] = f3{dy * r, y, -dx * r};
        for (int k = 0; k < 4; ++k) {
            f.eids[k] = kPfEdge | (uint32_t)(2 * k + 2);
            f.vids[k] = kPfVertex | (uint32_t)(2 * k + 1);
        }
        f.fid = kPfFace | 9u;
        f.n = 4;
    }
}

// ConvexPolyhedron's feature map (convex_polyhedron.rs:959-992): the face whose
// normal has the largest dot product with dir (the first on ties), its first four
// vertices and edges.
__device__ void polyhedron_feature(const ShapeRef &g, const pq_polyhedron_face *faces, const uint32_t *face_verts,
                                   const uint32_t *face_edges, f3 d, Feature &f) {
    const uint32_t first = __float_as_uint(g.p0), count = __float_as_uint(g.p1);
    uint32_t best = 0;
    float best_dot = dot3(f3{faces[first].normal[0], faces[first].normal[1], faces[first].normal[2]}, d);
    for (uint32_t k = 1; k < count; ++k) {
        const pq_polyhedron_face &F = faces[first + k];
        const float nd = dot3(f3{F.normal[0], F.normal[1], F.normal[2]}, d);
        if (nd > best_dot) {
            best = k;
            best_dot = nd;
        }
    }
    const pq_polyhedron_face F = faces[first + best];
    const int nv = F.count < 4u ? (int)F.count : 4;
    for (int k = 0; k < 4; ++k) {
        if (k < nv) {
            const uint32_t vid = face_verts[F.first + k];
            const float4 p = g.v[vid];
            f.v[k] = f3{p.x, p.y, p.z};
            f.vids[k] = kPfVertex | vid;
            f.eids[k] = kPfEdge | face_edges[F.first + k];
        } else {  // left as PolygonalFeature::default() has them
            f.v[k] = f3{0.0f, 0.0f, 0.0f};
            f.vids[k] = 0u;
            f.eids[k] = 0u;
        }
    }
    f.fid = kPfFace | best;
    f.n = nv;
}

__device__ void support_feature(const ShapeRef &g, const DeviceShapes &s, f3 d, Feature &f) {
    if (g.type == PQ_SHAPE_CONVEX_POLYHEDRON) polyhedron_feature(g, s.faces, s.face_verts, s.face_edges, d, f);
    else if (g.type == PQ_SHAPE_CUBOID) cuboid_feature(f3{g.p0, g.p1, g.p2}, d, f);
    else if (g.type == PQ_SHAPE_CONE) cone_feature(g.p0, g.p1, d, f);
    else if (g.type == PQ_SHAPE_TRIANGLE) triangle_feature(g, f);
    else segment_feature(g.p0, f);
}

// WBasis::orthonormal_basis (utils/wops.rs:96-109).
__device__ __forceinline__ void orthonormal_basis(f3 n, f3 &b0, f3 &b1) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    b0 = f3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    b1 = f3{b, sign + n.y * n.y * a, -n.y};
}

__device__ __forceinline__ P2 project2(f3 p, f3 b0, f3 b1) { return P2{dot3(p, b0), dot3(p, b1)}; }
__device__ __forceinline__ float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
__device__ __forceinline__ f3 inv_point(const iso &m, f3 p) { return qrot_inv(m.q, sub3(p, m.t)); }

struct Out {
    pq_contact_manifold *m;
    uint32_t n;
    __device__ void push(f3 p1, f3 p2, uint32_t fid1, uint32_t fid2, float dist) {
        if (n >= PQ_MANIFOLD_MAX_POINTS) return;  // cannot happen: at most 4 + 4 + 16 + 1 points
        pq_tracked_contact &c = m->points[n++];
        c.local_p1[0] = p1.x; c.local_p1[1] = p1.y; c.local_p1[2] = p1.z;
        c.local_p2[0] = p2.x; c.local_p2[1] = p2.y; c.local_p2[2] = p2.z;
        c.dist = dist;
        c.fid1 = fid1;
        c.fid2 = fid2;
    }
};

// closest_points_segment_segment_with_locations_nD in 2D (closest_points_segment_segment.rs:41-111):
// the barycentric coordinates of both closest points.
__device__ void seg_seg_2d(P2 a0, P2 a1, P2 b0, P2 b1, float bc1[2], float bc2[2]) {
    const P2 d1 = sub2(a1, a0), d2 = sub2(b1, b0), r = sub2(a0, b0);
    const float a = dot2(d1, d1), e = dot2(d2, d2), f = dot2(d2, r);
    float s, t;
    if (a <= kEps && e <= kEps) {
        s = 0.0f;
        t = 0.0f;
    } else if (a <= kEps) {
        s = 0.0f;
        t = clamp01(f / e);
    } else {
        const float c = dot2(d1, r);
        if (e <= kEps) {
            t = 0.0f;
            s = clamp01(-c / a);
        } else {
            const float b = dot2(d1, d2);
            const float ae = a * e, bb = b * b, denom = ae - bb;
            s = (denom > kEps && !ulps_eq(ae, bb)) ? clamp01((b * f - c * e) / denom) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = clamp01(-c / a);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = clamp01((b - c) / a);
            }
        }
    }
    if (s == 0.0f) { bc1[0] = 1.0f; bc1[1] = 0.0f; }
    else if (s == 1.0f) { bc1[0] = 0.0f; bc1[1] = 1.0f; }
    else { bc1[0] = 1.0f - s; bc1[1] = s; }
    if (t == 0.0f) { bc2[0] = 1.0f; bc2[1] = 0.0f; }
    else if (t == 1.0f) { bc2[0] = 0.0f; bc2[1] = 1.0f; }
    else { bc2[0] = 1.0f - t; bc2[1] = t; }
}

// PolygonalFeature::contacts_edge_edge (polygonal_feature3d.rs:94-200) with
// clip_segment_segment (clip_segment_segment.rs:76-140) for parallel edges.
__device__ void contacts_edge_edge(const iso &pos12, const Feature &f1, f3 sep, const Feature &f2, Out &o) {
    f3 b0, b1;
    orthonormal_basis(sep, b0, b1);
    const P2 pe10 = project2(f1.v[0], b0, b1), pe11 = project2(f1.v[1], b0, b1);
    const f3 v20 = iso_point(pos12, f2.v[0]), v21 = iso_point(pos12, f2.v[1]);
    const P2 pe20 = project2(v20, b0, b1), pe21 = project2(v21, b0, b1);
    const P2 u1 = sub2(pe11, pe10), u2 = sub2(pe21, pe20);
    const float n1 = sqrtf(u1.x * u1.x + u1.y * u1.y), n2 = sqrtf(u2.x * u2.x + u2.y * u2.y);
    if (n1 > kEps && n2 > kEps) {
        const P2 t1{u1.x / n1, u1.y / n1}, t2{u2.x / n2, u2.y / n2};
        if (!(dot2(t1, t2) >= kCosFracPi8)) {
            float bc1[2], bc2[2];
            seg_seg_2d(pe10, pe11, pe20, pe21, bc1, bc2);
            const f3 p1 = add3(mul3(f1.v[0], bc1[0]), mul3(f1.v[1], bc1[1]));
            const f3 p21 = add3(mul3(v20, bc2[0]), mul3(v21, bc2[1]));
            o.push(p1, inv_point(pos12, p21), f1.eids[0], f2.eids[0], dot3(sub3(p21, p1), sep));
            return;
        }
    }
    // clip_segment_segment((f1.v0, f1.v1), (v20, v21))
    f3 s10 = f1.v[0], s11 = f1.v[1], s20 = v20, s21 = v21;
    const f3 tangent1 = sub3(s11, s10);
    float r1a = 0.0f, r1b = norm_sq3(tangent1);
    float r2a = dot3(sub3(s20, s10), tangent1), r2b = dot3(sub3(s21, s10), tangent1);
    int fa1 = 0, fb1 = 2, fa2 = 0, fb2 = 2;
    if (r1b < r1a) {
        const float t = r1a; r1a = r1b; r1b = t;
        fa1 = 2; fb1 = 0;
        const f3 q = s10; s10 = s11; s11 = q;
    }
    if (r2b < r2a) {
        const float t = r2a; r2a = r2b; r2b = t;
        fa2 = 2; fb2 = 0;
        const f3 q = s20; s20 = s21; s21 = q;
    }
    if (r2a > r1b || r1a > r2b) return;
    const float length1 = r1b - r1a, length2 = r2b - r2a;
    f3 ca1, ca2, cb1, cb2;
    int caf1, caf2, cbf1, cbf2;
    if (r2a > r1a) {
        ca1 = add3(s10, mul3(tangent1, (r2a - r1a) / length1));
        ca2 = s20;
        caf1 = 1;
        caf2 = fa2;
    } else {
        ca1 = s10;
        ca2 = add3(s20, mul3(sub3(s21, s20), (r1a - r2a) / length2));
        caf1 = fa1;
        caf2 = 1;
    }
    if (r2b < r1b) {
        cb1 = add3(s10, mul3(tangent1, (r2b - r1a) / length1));
        cb2 = s21;
        cbf1 = 1;
        cbf2 = fb2;
    } else {
        cb1 = s11;
        cb2 = add3(s20, mul3(sub3(s21, s20), (r1b - r2a) / length2));
        cbf1 = fb1;
        cbf2 = 1;
    }
    auto feat = [](const Feature &f, int id) { return id == 0 ? f.vids[0] : (id == 1 ? f.eids[0] : f.vids[1]); };
    o.push(ca1, inv_point(pos12, ca2), feat(f1, caf1), feat(f2, caf2), dot3(sub3(ca2, ca1), sep));
    o.push(cb1, inv_point(pos12, cb2), feat(f1, cbf1), feat(f2, cbf2), dot3(sub3(cb2, cb1), sep));
}

// closest_points_line2d (polygonal_feature3d.rs:383-439).
__device__ bool closest_points_line2d(P2 e10, P2 e11, P2 e20, P2 e21, float &s_out, float &t_out) {
    const P2 dir1 = sub2(e11, e10), dir2 = sub2(e21, e20), r = sub2(e10, e20);
    const float a = dot2(dir1, dir1), e = dot2(dir2, dir2), f = dot2(dir2, r);
    if (a <= kEps && e <= kEps) {
        s_out = 0.0f;
        t_out = 0.0f;
        return true;
    }
    if (a <= kEps) {
        s_out = 0.0f;
        t_out = f / e;
        return true;
    }
    const float c = dot2(dir1, r);
    if (e <= kEps) {
        s_out = -c / a;
        t_out = 0.0f;
        return true;
    }
    const float b = dot2(dir1, dir2);
    const float ae = a * e, bb = b * b, denom = ae - bb;
    if (denom <= kEps || ulps_eq(ae, bb)) return false;
    const float s = (b * f - c * e) / denom;
    s_out = s;
    t_out = (b * s + f) / e;
    return true;
}

// Whether the projected point p lies inside the projected polygon pf[0..n)
// (the sign test of polygonal_feature3d.rs:265-279).
__device__ __forceinline__ bool inside_polygon(const P2 *pf, int n, P2 p) {
    const int last = n - 1;
    float sign = perp2(sub2(pf[0], pf[last]), sub2(p, pf[last]));
    for (int j = 0; j < last; ++j) {
        const float ns = perp2(sub2(pf[j + 1], pf[j]), sub2(p, pf[j]));
        if (sign == 0.0f) sign = ns;
        else if (sign * ns < 0.0f) return false;
    }
    return true;
}

// PolygonalFeature::contacts_face_face (polygonal_feature3d.rs:202-381).
__device__ void contacts_face_face(const iso &pos12, const Feature &f1, f3 sep, const Feature &f2, Out &o) {
    f3 b0, b1;
    orthonormal_basis(sep, b0, b1);
    P2 pf1[4], pf2[4];
    f3 v21[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pf1[k] = project2(f1.v[k], b0, b1);
        v21[k] = iso_point(pos12, f2.v[k]);
        pf2[k] = project2(v21[k], b0, b1);
    }
    if (f2.n > 2) {
        const f3 normal2_1 = cross3(sub3(v21[2], v21[1]), sub3(v21[0], v21[1]));
        const float denom = dot3(normal2_1, sep);
        if (!rel_eq1(denom, 0.0f)) {
            for (int i = 0; i < f1.n; ++i) {
                if (!inside_polygon(pf2, f2.n, pf1[i])) continue;
                const float dist = dot3(sub3(v21[0], f1.v[i]), normal2_1) / denom;
                const f3 lp21 = add3(f1.v[i], mul3(sep, dist));
                o.push(f1.v[i], inv_point(pos12, lp21), f1.vids[i], f2.fid, dist);
            }
        }
    }
    if (f1.n > 2) {
        const f3 normal1 = cross3(sub3(f1.v[2], f1.v[1]), sub3(f1.v[0], f1.v[1]));
        const float denom = -dot3(normal1, sep);
        if (!rel_eq1(denom, 0.0f)) {
            for (int i = 0; i < f2.n; ++i) {
                if (!inside_polygon(pf1, f1.n, pf2[i])) continue;
                const float dist = dot3(sub3(f1.v[0], v21[i]), normal1) / denom;
                const f3 lp1 = sub3(v21[i], mul3(sep, dist));
                o.push(lp1, inv_point(pos12, v21[i]), f1.fid, f2.vids[i], dist);
            }
        }
    }
    for (int j = 0; j < f2.n; ++j) {
        const int j1 = (j + 1) % f2.n;
        for (int i = 0; i < f1.n; ++i) {
            const int i1 = (i + 1) % f1.n;
            float s, t;
            if (!closest_points_line2d(pf1[i], pf1[i1], pf2[j], pf2[j1], s, t)) continue;
            if (s > 0.0f && s < 1.0f && t > 0.0f && t < 1.0f) {
                const f3 lp1 = add3(mul3(f1.v[i], 1.0f - s), mul3(f1.v[i1], s));
                const f3 lp21 = add3(mul3(v21[j], 1.0f - t), mul3(v21[j1], t));
                o.push(lp1, inv_point(pos12, lp21), f1.eids[i], f2.eids[j], dot3(sub3(lp21, lp1), sep));
            }
        }
    }
}

// The manifold of pair k of `pairs` (its GJK/EPA contact in contacts[k]) goes to
// out[list[k]], or out[k] without a list; m_dev, if given, is the device-side
// number of pairs.  keep_n2: out holds last frame's manifolds, whose local_n2 a
// pair without ClosestPoints keeps (ContactManifold::clear drops only the
// points, contact_manifolds_pfm_pfm.rs:150-157); a fresh manifold has it zero.
__global__ __launch_bounds__(kBlock) void manifold_kernel(const DeviceShapes s, const float *__restrict__ poses,
                                                          const uint32_t *__restrict__ pairs, uint64_t m,
                                                          const unsigned long long *__restrict__ m_dev,
                                                          const uint32_t *__restrict__ list,
                                                          const pq_contact *__restrict__ contacts,
                                                          pq_contact_manifold *__restrict__ out, int keep_n2) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= (m_dev ? (uint64_t)*m_dev : m)) return;
    pq_contact_manifold *const mf = out + (list ? list[k] : k);
    const pq_contact c = contacts[k];
    Out o{mf, 0u};
    f3 n1{0.0f, 0.0f, 0.0f}, n2{0.0f, 0.0f, 0.0f};
    if (c.has_contact == 1u) {
        const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const iso pos12 = iso_inv_mul(load_iso(poses + 7 * (size_t)i), load_iso(poses + 7 * (size_t)j));
        ShapeRef g1 = load_shape(s.rec, s.verts, i), g2 = load_shape(s.rec, s.verts, j);
        float br1 = 0.0f, br2 = 0.0f;
        if (g1.type == PQ_SHAPE_CAPSULE) {
            g1.type = kShapeSegment;
            br1 = g1.p1;
        }
        if (g2.type == PQ_SHAPE_CAPSULE) {
            g2.type = kShapeSegment;
            br2 = g2.p1;
        }
        n1 = f3{c.normal1[0], c.normal1[1], c.normal1[2]};
        n2 = f3{c.normal2[0], c.normal2[1], c.normal2[2]};
        Feature f1, f2;
        support_feature(g1, s, n1, f1);
        support_feature(g2, s, n2, f2);
        if (f1.n == 2 && f2.n == 2) contacts_edge_edge(pos12, f1, n1, f2, o);
        else contacts_face_face(pos12, f1, n1, f2, o);
        // GJK's own point (contact_manifolds_pfm_pfm.rs:108-120)
        o.push(f3{c.point1[0], c.point1[1], c.point1[2]}, f3{c.point2[0], c.point2[1], c.point2[2]}, 0u, 0u, c.dist);
        if (br1 != 0.0f || br2 != 0.0f) {  // contact_manifolds_pfm_pfm.rs:134-141
            for (uint32_t q = 0; q < o.n; ++q) {
                pq_tracked_contact &t = mf->points[q];
                t.local_p1[0] += n1.x * br1; t.local_p1[1] += n1.y * br1; t.local_p1[2] += n1.z * br1;
                t.local_p2[0] += n2.x * br2; t.local_p2[1] += n2.y * br2; t.local_p2[2] += n2.z * br2;
                t.dist -= br1 + br2;
            }
        }
        mf->status = 1u;
    } else {
        // NoIntersection(dir): the cached normal; anything else: unsupported
        n1 = f3{c.normal1[0], c.normal1[1], c.normal1[2]};
        mf->status = c.has_contact;
    }
    mf->local_n1[0] = n1.x; mf->local_n1[1] = n1.y; mf->local_n1[2] = n1.z;
    if (c.has_contact == 1u || !keep_n2) {
        mf->local_n2[0] = n2.x; mf->local_n2[1] = n2.y; mf->local_n2[2] = n2.z;
    }
    mf->num_points = o.n;
}

// ContactManifold::try_update_contacts_eps (contact_manifold.rs:662-699) with the
// thresholds contact_manifold_pfm_pfm gives it (COS_1_DEGREES and 1e-6,
// contact_manifolds_pfm_pfm.rs:60-65), one lane per pair.  A manifold it keeps is
// updated in place and marked PQ_MANIFOLD_UPDATED.  Every other pair goes on the
// list the contact and manifold kernels run on, with its GJK seed: the cached
// normal as Unit::try_new(local_n1, DEFAULT_EPSILON) makes it (zero, for parry's
// default direction, when that is None).  parry writes each point as it checks
// it and drops them all when one fails; checking them all before writing any
// leaves the same manifold.
__global__ __launch_bounds__(kBlock) void manifold_update_kernel(const float *__restrict__ poses,
                                                                 const uint32_t *__restrict__ pairs, uint64_t m,
                                                                 pq_contact_manifold *__restrict__ mf,
                                                                 uint32_t *__restrict__ list,
                                                                 uint32_t *__restrict__ sub_pairs,
                                                                 float *__restrict__ seeds,
                                                                 unsigned long long *__restrict__ pending) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool pend = false;
    uint32_t i = 0, j = 0;
    f3 seed{0.0f, 0.0f, 0.0f};
    if (k < m) {
        i = pairs[2 * k];
        j = pairs[2 * k + 1];
        pq_contact_manifold *const M = mf + k;
        const uint32_t np = M->num_points;
        const f3 n1{M->local_n1[0], M->local_n1[1], M->local_n1[2]};
        bool kept = np != 0u && np <= (uint32_t)PQ_MANIFOLD_MAX_POINTS;
        if (kept) {
            const iso pos12 = iso_inv_mul(load_iso(poses + 7 * (size_t)i), load_iso(poses + 7 * (size_t)j));
            const f3 ln2 = qrot(pos12.q, f3{M->local_n2[0], M->local_n2[1], M->local_n2[2]});
            kept = !(-dot3(n1, ln2) < kCos1Degrees);
            for (uint32_t q = 0; kept && q < np; ++q) {
                const pq_tracked_contact &t = M->points[q];
                const f3 p1{t.local_p1[0], t.local_p1[1], t.local_p1[2]};
                const f3 lp2 = iso_point(pos12, f3{t.local_p2[0], t.local_p2[1], t.local_p2[2]});
                const float dist = dot3(sub3(lp2, p1), n1);
                if (dist * t.dist < 0.0f) kept = false;  // between penetrating and not
                else if (norm_sq3(sub3(sub3(lp2, mul3(n1, dist)), p1)) > 1.0e-6f) kept = false;
            }
            if (kept) {
                for (uint32_t q = 0; q < np; ++q) {
                    pq_tracked_contact &t = M->points[q];
                    const f3 p1{t.local_p1[0], t.local_p1[1], t.local_p1[2]};
                    const f3 lp2 = iso_point(pos12, f3{t.local_p2[0], t.local_p2[1], t.local_p2[2]});
                    const float dist = dot3(sub3(lp2, p1), n1);
                    const f3 np1 = sub3(lp2, mul3(n1, dist));
                    t.dist = dist;
                    t.local_p1[0] = np1.x;
                    t.local_p1[1] = np1.y;
                    t.local_p1[2] = np1.z;
                }
                M->status = PQ_MANIFOLD_UPDATED;
            }
        }
        if (!kept) {
            pend = true;
            const float sq = norm_sq3(n1);
            if (sq > kEps * kEps) seed = div3(n1, sqrtf(sq));
        }
    }
    // Wave-aggregated append of the pairs to recompute.
    const uint64_t mask = __ballot(pend);
    if (mask) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(pending, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        if (pend) {
            const uint64_t p = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            list[p] = (uint32_t)k;
            sub_pairs[2 * p] = i;
            sub_pairs[2 * p + 1] = j;
            seeds[3 * p] = seed.x;
            seeds[3 * p + 1] = seed.y;
            seeds[3 * p + 2] = seed.z;
        }
    }
}

__global__ void zero_pending_kernel(unsigned long long *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

struct ManifoldWorkspace {
    pq_contact *contacts;
    void *contact_ws;
    uint32_t *list, *sub_pairs;
    float *seeds;
    unsigned long long *pending;
    size_t bytes;
};

ManifoldWorkspace carve(void *base, uint64_t m) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    char *ws = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = ws ? ws + off : nullptr;
        off += up(bytes);
        return p;
    };
    ManifoldWorkspace w{};
    w.contacts = (pq_contact *)take(m * sizeof(pq_contact));
    w.contact_ws = take(contact_workspace_bytes(m));
    w.list = (uint32_t *)take(m * sizeof(uint32_t));
    w.sub_pairs = (uint32_t *)take(2 * m * sizeof(uint32_t));
    w.seeds = (float *)take(3 * m * sizeof(float));
    w.pending = (unsigned long long *)take(sizeof(unsigned long long));
    w.bytes = off;
    return w;
}

}  // namespace

size_t manifold_workspace_bytes(uint64_t m) { return carve(nullptr, m ? m : 1).bytes; }

hipError_t launch_manifolds(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                            float prediction, pq_contact_manifold *out, void *workspace, hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const ManifoldWorkspace w = carve(workspace, m);
    hipError_t e = launch_contacts(s, poses, pairs, m, nullptr, prediction, nullptr, w.contacts, nullptr, w.contact_ws,
                                   stream, kContactFeatureMaps);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(manifold_kernel, dim3((uint32_t)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, s, poses,
                       pairs, m, nullptr, nullptr, w.contacts, out, 0);
    return hipGetLastError();
}

hipError_t launch_manifolds_update(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                                   float prediction, pq_contact_manifold *manifolds, void *workspace,
                                   hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const ManifoldWorkspace w = carve(workspace, m);
    const dim3 grid((uint32_t)((m + kBlock - 1) / kBlock)), block(kBlock);
    hipLaunchKernelGGL(zero_pending_kernel, dim3(1), dim3(64), 0, stream, w.pending);
    hipLaunchKernelGGL(manifold_update_kernel, grid, block, 0, stream, poses, pairs, m, manifolds, w.list, w.sub_pairs,
                       w.seeds, w.pending);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_contacts(s, poses, w.sub_pairs, m, (const uint64_t *)w.pending, prediction, w.seeds, w.contacts,
                        nullptr, w.contact_ws, stream, kContactFeatureMaps);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(manifold_kernel, grid, block, 0, stream, s, poses, w.sub_pairs, m, w.pending, w.list,
                       w.contacts, manifolds, 1);
    return hipGetLastError();
}

}  // namespace pq
