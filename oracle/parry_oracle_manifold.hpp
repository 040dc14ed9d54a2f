// parry_oracle_manifold.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of parry3d's contact manifold between two polygonal feature
// maps, contact_manifold_pfm_pfm (src/query/contact_manifolds/
// contact_manifolds_pfm_pfm.rs:44-162), on a fresh manifold or on the previous
// frame's (try_update_contacts_eps, contact_manifold.rs:662-699, then GJK seeded
// by the cached normal), with no normal constraints; match_contacts moves user
// data only and is not modelled.  The feature maps (polygonal_feature_map.rs:30-154,
// cuboid.rs:267-352, triangle.rs:280-282) and the clipping
// (polygonal_feature3d.rs:74-439: contacts_edge_edge, contacts_face_face,
// closest_points_line2d; clip_segment_segment.rs:76-140;
// closest_points_segment_segment.rs:41-111; the orthonormal basis of
// utils/wops.rs:92-110) follow the reference's operation order.  Same
// conventions as parry_oracle.hpp: T = float is parry3d, T = double parry3d-f64.
// A Capsule's feature map is its segment with the capsule radius as border
// (shape.rs:855-857); a Ball has none; a ConvexPolyhedron's uses the faces given
// with it (convex_polyhedron.rs:959-992).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "parry_oracle.hpp"
#include "parry_oracle_contact.hpp"

namespace oracle {

// PackedFeatureId (feature_id.rs:446-543).
constexpr uint32_t PF_UNKNOWN = 0u;
inline uint32_t pf_vertex(uint32_t c) { return (1u << 30) | c; }
inline uint32_t pf_edge(uint32_t c) { return (2u << 30) | c; }
inline uint32_t pf_face(uint32_t c) { return (3u << 30) | c; }

template <class T> struct PolyFeature {
    V3<T> v[4];
    uint32_t vids[4], eids[4], fid;
    int n;
};

template <class T> struct TrackedContact {
    V3<T> p1, p2;
    uint32_t fid1, fid2;
    T dist;
};

template <class T> struct Manifold {
    V3<T> n1, n2;
    std::vector<TrackedContact<T>> points;
};

template <class T> constexpr T cos_frac_pi_8() { return T(0.92387953251); }  // utils/consts.rs:11

// approx::ulps_eq!(a, b) with epsilon = EPS and max_ulps = 4.
inline bool ulps_eq(float a, float b) {
    float d = a > b ? a - b : b - a;
    if (d <= Eps<float>::eps) return true;
    if (std::signbit(a) != std::signbit(b) || std::isnan(a) || std::isnan(b)) return false;
    uint32_t ia, ib;
    std::memcpy(&ia, &a, 4);
    std::memcpy(&ib, &b, 4);
    return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}
inline bool ulps_eq(double a, double b) {
    double d = a > b ? a - b : b - a;
    if (d <= Eps<double>::eps) return true;
    if (std::signbit(a) != std::signbit(b) || std::isnan(a) || std::isnan(b)) return false;
    uint64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    return (ia <= ib ? ib - ia : ia - ib) <= 4u;
}

template <class T> struct P2 {
    T x, y;
};
template <class T> inline P2<T> sub2(P2<T> a, P2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline T dot2(P2<T> a, P2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T perp2(P2<T> a, P2<T> b) { return a.x * b.y - a.y * b.x; }

// ---- feature maps -------------------------------------------------------------

// Segment a = (0, -h, 0), b = (0, h, 0) (polygonal_feature3d.rs:43-53).
template <class T> PolyFeature<T> segment_feature(T h) {
    PolyFeature<T> f{};
    V3<T> a = v3(T(0), -h, T(0)), b = v3(T(0), h, T(0));
    f.v[0] = a;
    f.v[1] = b;
    f.v[2] = b;
    f.v[3] = b;
    const uint32_t vi[4] = {0, 1, 1, 1};
    for (int k = 0; k < 4; ++k) {
        f.vids[k] = pf_vertex(vi[k]);
        f.eids[k] = pf_edge(0);
    }
    f.fid = pf_face(0);
    f.n = 2;
    return f;
}

// Triangle::support_face = the whole triangle (triangle.rs:280-282, polygonal_feature3d.rs:31-41).
template <class T> PolyFeature<T> triangle_feature(const V3<T> *p) {
    PolyFeature<T> f{};
    f.v[0] = p[0];
    f.v[1] = p[1];
    f.v[2] = p[2];
    f.v[3] = p[2];
    const uint32_t ids[4] = {0, 1, 2, 2};
    for (int k = 0; k < 4; ++k) {
        f.vids[k] = pf_vertex(ids[k]);
        f.eids[k] = pf_edge(ids[k]);
    }
    f.fid = pf_face(0);
    f.n = 3;
    return f;
}

// nalgebra iamax: the first index of the largest |component|.
template <class T> inline int iamax3(V3<T> d) {
    int i = 0;
    T m = std::fabs(d.x);
    if (std::fabs(d.y) > m) {
        m = std::fabs(d.y);
        i = 1;
    }
    if (std::fabs(d.z) > m) i = 2;
    return i;
}

// Cuboid::support_face (cuboid.rs:267-352).
template <class T> PolyFeature<T> cuboid_feature(V3<T> he, V3<T> d) {
    PolyFeature<T> f{};
    const int ia = iamax3(d);
    const T s = std::copysign(T(1), d[ia]);
    if (ia == 0) {
        f.v[0] = v3(he.x * s, he.y, he.z);
        f.v[1] = v3(he.x * s, -he.y, he.z);
        f.v[2] = v3(he.x * s, -he.y, -he.z);
        f.v[3] = v3(he.x * s, he.y, -he.z);
    } else if (ia == 1) {
        f.v[0] = v3(he.x, he.y * s, he.z);
        f.v[1] = v3(-he.x, he.y * s, he.z);
        f.v[2] = v3(-he.x, he.y * s, -he.z);
        f.v[3] = v3(he.x, he.y * s, -he.z);
    } else {
        f.v[0] = v3(he.x, he.y, he.z * s);
        f.v[1] = v3(he.x, -he.y, he.z * s);
        f.v[2] = v3(-he.x, -he.y, he.z * s);
        f.v[3] = v3(-he.x, he.y, he.z * s);
    }
    const int si = ((int)(int8_t)s + 1) / 2;
    static const uint32_t VIDS[3][2][4] = {{{0, 4, 6, 2}, {8, 12, 14, 10}},
                                           {{0, 8, 10, 2}, {4, 12, 14, 6}},
                                           {{0, 4, 12, 8}, {2, 6, 14, 10}}};
    static const uint32_t EIDS[3][2][4] = {
        {{0b11010000, 0b11011010, 0b11011001, 0b11001000}, {0b11110100, 0b11111110, 0b11111101, 0b11101100}},
        {{0b11100000, 0b11101100, 0b11101001, 0b11001000}, {0b11110010, 0b11111110, 0b11111011, 0b11011010}},
        {{0b11010000, 0b11110010, 0b11110100, 0b11100000}, {0b11011001, 0b11111011, 0b11111101, 0b11101001}}};
    for (int k = 0; k < 4; ++k) {
        f.vids[k] = pf_vertex(VIDS[ia][si][k]);
        f.eids[k] = pf_edge(EIDS[ia][si][k]);
    }
    f.fid = pf_face((uint32_t)(ia + si * 3 + 10));
    f.n = 4;
    return f;
}

// Cone's feature map (polygonal_feature_map.rs:108-154); h = half height, r = radius.
template <class T> PolyFeature<T> cone_feature(T h, T r, V3<T> d) {
    PolyFeature<T> f{};
    for (int k = 0; k < 4; ++k) f.v[k] = v3(T(0), T(0), T(0));
    T dx = T(1), dy = T(0);
    {
        T nrm = std::sqrt(d.x * d.x + d.z * d.z);  // Vector2::try_normalize(EPS)
        if (nrm > Eps<T>::eps) {
            dx = d.x / nrm;
            dy = d.z / nrm;
        }
    }
    if (d.y > T(0)) {
        f.v[0] = v3(dx * r, -h, dy * r);
        f.v[1] = v3(T(0), h, T(0));
        const uint32_t vi[4] = {1, 11, 11, 11};
        for (int k = 0; k < 4; ++k) {
            f.eids[k] = pf_edge(0);
            f.vids[k] = pf_vertex(vi[k]);
        }
        f.fid = pf_face(0);
        f.n = 2;
    } else {
        const T y = -h;
        f.v[0] = v3(dx * r, y, dy * r);
        f.v[1] = v3(-dy * r, y, dx * r);
        f.v[2] = v3(-dx * r, y, -dy * r);
        f.v[3] = v3(dy * r, y, -dx * r);
        const uint32_t ei[4] = {2, 4, 6, 8}, vi[4] = {1, 3, 5, 7};
        for (int k = 0; k < 4; ++k) {
            f.eids[k] = pf_edge(ei[k]);
            f.vids[k] = pf_vertex(vi[k]);
        }
        f.fid = pf_face(9);
        f.n = 4;
    }
    return f;
}

// ConvexPolyhedron's feature map (convex_polyhedron.rs:959-992): the first face
// with the largest normal . dir, its first four vertices and edges.
template <class T> PolyFeature<T> polyhedron_feature(const Shape<T> &s, V3<T> d) {
    PolyFeature<T> f{};
    uint32_t best = 0;
    T best_dot = dot(s.faces[0].normal, d);
    for (uint32_t k = 1; k < s.nfaces; ++k) {
        T nd = dot(s.faces[k].normal, d);
        if (nd > best_dot) {
            best = k;
            best_dot = nd;
        }
    }
    const PolyFace<T> &F = s.faces[best];
    const int nv = F.count < 4u ? (int)F.count : 4;
    for (int k = 0; k < 4; ++k) {
        if (k < nv) {
            const uint32_t vid = s.fverts[F.first + k];
            f.v[k] = s.pts[vid];
            f.vids[k] = pf_vertex(vid);
            f.eids[k] = pf_edge(s.fedges[F.first + k]);
        } else {
            f.v[k] = v3(T(0), T(0), T(0));
            f.vids[k] = f.eids[k] = PF_UNKNOWN;
        }
    }
    f.fid = pf_face(best);
    f.n = nv;
    return f;
}

// The feature map and border radius of a shape (shape.rs as_polygonal_feature_map).
// Returns false for shapes without one here.
template <class T> bool pfm_of(const Shape<T> &s, Shape<T> &pfm, T &border) {
    border = T(0);
    pfm = s;
    switch (s.type) {
        case CUBOID:
        case CONE:
        case TRIANGLE:
            return true;
        case CONVEX_POLYHEDRON:
            return s.nfaces > 0;
        case CAPSULE:
            pfm.type = SEGMENT;
            border = s.p[1];
            return true;
        default:
            return false;
    }
}

template <class T> PolyFeature<T> support_feature(const Shape<T> &pfm, V3<T> d) {
    switch (pfm.type) {
        case CONVEX_POLYHEDRON: return polyhedron_feature(pfm, d);
        case CUBOID: return cuboid_feature(v3(pfm.p[0], pfm.p[1], pfm.p[2]), d);
        case CONE: return cone_feature(pfm.p[0], pfm.p[1], d);
        case TRIANGLE: return triangle_feature(pfm.pts);
        default: return segment_feature(pfm.p[0]);
    }
}

// ---- clipping ---------------------------------------------------------------------

// WBasis::orthonormal_basis (utils/wops.rs:96-109).
template <class T> void orthonormal_basis(V3<T> n, V3<T> &b0, V3<T> &b1) {
    const T sign = std::copysign(T(1), n.z);
    const T a = -T(1) / (sign + n.z);
    const T b = n.x * n.y * a;
    b0 = v3(T(1) + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b1 = v3(b, sign + n.y * n.y * a, -n.y);
}

template <class T> inline V3<T> iso_apply(const Iso<T> &m, V3<T> p) { return transform_point(m, p); }
template <class T> inline V3<T> iso_inv_point(const Iso<T> &m, V3<T> p) { return inv_rotate(m.q, p - m.t); }

template <class T> inline T clamp01(T x) { return x < T(0) ? T(0) : (x > T(1) ? T(1) : x); }

// closest_points_segment_segment_with_locations_nD (closest_points_segment_segment.rs:41-111)
// in 2D; returns the barycentric coordinates of both locations.
template <class T> void seg_seg_2d(P2<T> a0, P2<T> a1, P2<T> b0, P2<T> b1, T bc1[2], T bc2[2]) {
    P2<T> d1 = sub2(a1, a0), d2 = sub2(b1, b0), r = sub2(a0, b0);
    T a = dot2(d1, d1), e = dot2(d2, d2), f = dot2(d2, r);
    T s, t;
    const T eps = Eps<T>::eps;
    if (a <= eps && e <= eps) {
        s = T(0);
        t = T(0);
    } else if (a <= eps) {
        s = T(0);
        t = clamp01(f / e);
    } else {
        T c = dot2(d1, r);
        if (e <= eps) {
            t = T(0);
            s = clamp01(-c / a);
        } else {
            T b = dot2(d1, d2);
            T ae = a * e, bb = b * b, denom = ae - bb;
            if (denom > eps && !ulps_eq(ae, bb)) s = clamp01((b * f - c * e) / denom);
            else s = T(0);
            t = (b * s + f) / e;
            if (t < T(0)) {
                t = T(0);
                s = clamp01(-c / a);
            } else if (t > T(1)) {
                t = T(1);
                s = clamp01((b - c) / a);
            }
        }
    }
    // SegmentPointLocation::barycentric_coordinates: OnVertex(0) = [1, 0], OnVertex(1) = [0, 1]
    if (s == T(0)) { bc1[0] = T(1); bc1[1] = T(0); }
    else if (s == T(1)) { bc1[0] = T(0); bc1[1] = T(1); }
    else { bc1[0] = T(1) - s; bc1[1] = s; }
    if (t == T(0)) { bc2[0] = T(1); bc2[1] = T(0); }
    else if (t == T(1)) { bc2[0] = T(0); bc2[1] = T(1); }
    else { bc2[0] = T(1) - t; bc2[1] = t; }
}

template <class T> struct ClipPoint {
    V3<T> p1, p2;
    int f1, f2;
};

// clip_segment_segment (clip_segment_segment.rs:76-140).
template <class T> bool clip_segment_segment(V3<T> s10, V3<T> s11, V3<T> s20, V3<T> s21, ClipPoint<T> &ca,
                                             ClipPoint<T> &cb) {
    V3<T> tangent1 = s11 - s10;
    T sq = dot(tangent1, tangent1);
    T range1[2] = {T(0), sq};
    T range2[2] = {dot(s20 - s10, tangent1), dot(s21 - s10, tangent1)};
    int features1[2] = {0, 2}, features2[2] = {0, 2};
    if (range1[1] < range1[0]) {
        std::swap(range1[0], range1[1]);
        std::swap(features1[0], features1[1]);
        std::swap(s10, s11);
    }
    if (range2[1] < range2[0]) {
        std::swap(range2[0], range2[1]);
        std::swap(features2[0], features2[1]);
        std::swap(s20, s21);
    }
    if (range2[0] > range1[1] || range1[0] > range2[1]) return false;
    T length1 = range1[1] - range1[0];
    T length2 = range2[1] - range2[0];
    if (range2[0] > range1[0]) {
        T bcoord = (range2[0] - range1[0]) / length1;
        ca = ClipPoint<T>{s10 + tangent1 * bcoord, s20, 1, features2[0]};
    } else {
        T bcoord = (range1[0] - range2[0]) / length2;
        ca = ClipPoint<T>{s10, s20 + (s21 - s20) * bcoord, features1[0], 1};
    }
    if (range2[1] < range1[1]) {
        T bcoord = (range2[1] - range1[0]) / length1;
        cb = ClipPoint<T>{s10 + tangent1 * bcoord, s21, 1, features2[1]};
    } else {
        T bcoord = (range1[1] - range2[0]) / length2;
        cb = ClipPoint<T>{s11, s20 + (s21 - s20) * bcoord, features1[1], 1};
    }
    return true;
}

// closest_points_line2d (polygonal_feature3d.rs:383-439).
template <class T> bool closest_points_line2d(P2<T> e10, P2<T> e11, P2<T> e20, P2<T> e21, T &s_out, T &t_out) {
    P2<T> dir1 = sub2(e11, e10), dir2 = sub2(e21, e20), r = sub2(e10, e20);
    T a = dot2(dir1, dir1), e = dot2(dir2, dir2), f = dot2(dir2, r);
    const T eps = Eps<T>::eps;
    if (a <= eps && e <= eps) {
        s_out = T(0);
        t_out = T(0);
        return true;
    }
    if (a <= eps) {
        s_out = T(0);
        t_out = f / e;
        return true;
    }
    T c = dot2(dir1, r);
    if (e <= eps) {
        s_out = -c / a;
        t_out = T(0);
        return true;
    }
    T b = dot2(dir1, dir2);
    T ae = a * e, bb = b * b, denom = ae - bb;
    bool parallel = denom <= eps || ulps_eq(ae, bb);
    if (parallel) return false;
    T s = (b * f - c * e) / denom;
    s_out = s;
    t_out = (b * s + f) / e;
    return true;
}

template <class T> inline P2<T> project2(V3<T> p, V3<T> b0, V3<T> b1) { return P2<T>{dot(p, b0), dot(p, b1)}; }

// Vector2::try_normalize(EPS).
template <class T> inline bool try_normalize2(P2<T> v, P2<T> &u) {
    T n = std::sqrt(v.x * v.x + v.y * v.y);
    if (n <= Eps<T>::eps) return false;
    u = P2<T>{v.x / n, v.y / n};
    return true;
}

// PolygonalFeature::contacts_edge_edge (polygonal_feature3d.rs:94-200), flipped = false.
template <class T>
void contacts_edge_edge(const Iso<T> &pos12, const PolyFeature<T> &f1, V3<T> sep, const PolyFeature<T> &f2,
                        std::vector<TrackedContact<T>> &out) {
    V3<T> b0, b1;
    orthonormal_basis(sep, b0, b1);
    P2<T> pe1[2] = {project2(f1.v[0], b0, b1), project2(f1.v[1], b0, b1)};
    V3<T> v21[2] = {iso_apply(pos12, f2.v[0]), iso_apply(pos12, f2.v[1])};
    P2<T> pe2[2] = {project2(v21[0], b0, b1), project2(v21[1], b0, b1)};
    P2<T> t1, t2;
    bool h1 = try_normalize2(sub2(pe1[1], pe1[0]), t1);
    bool h2 = try_normalize2(sub2(pe2[1], pe2[0]), t2);
    if (h1 && h2) {
        bool parallel = dot2(t1, t2) >= cos_frac_pi_8<T>();
        if (!parallel) {
            T bc1[2], bc2[2];
            seg_seg_2d(pe1[0], pe1[1], pe2[0], pe2[1], bc1, bc2);
            V3<T> p1 = f1.v[0] * bc1[0] + f1.v[1] * bc1[1];
            V3<T> p21 = v21[0] * bc2[0] + v21[1] * bc2[1];
            T dist = dot(p21 - p1, sep);
            out.push_back(TrackedContact<T>{p1, iso_inv_point(pos12, p21), f1.eids[0], f2.eids[0], dist});
            return;
        }
    }
    ClipPoint<T> ca, cb;
    if (clip_segment_segment(f1.v[0], f1.v[1], v21[0], v21[1], ca, cb)) {
        auto feat = [](const PolyFeature<T> &f, int id) { return id == 0 ? f.vids[0] : (id == 1 ? f.eids[0] : f.vids[1]); };
        out.push_back(TrackedContact<T>{ca.p1, iso_inv_point(pos12, ca.p2), feat(f1, ca.f1), feat(f2, ca.f2),
                                        dot(ca.p2 - ca.p1, sep)});
        out.push_back(TrackedContact<T>{cb.p1, iso_inv_point(pos12, cb.p2), feat(f1, cb.f1), feat(f2, cb.f2),
                                        dot(cb.p2 - cb.p1, sep)});
    }
}

// PolygonalFeature::contacts_face_face (polygonal_feature3d.rs:202-381), flipped = false.
template <class T>
void contacts_face_face(const Iso<T> &pos12, const PolyFeature<T> &f1, V3<T> sep, const PolyFeature<T> &f2,
                        std::vector<TrackedContact<T>> &out) {
    V3<T> b0, b1;
    orthonormal_basis(sep, b0, b1);
    P2<T> pf1[4], pf2[4];
    V3<T> v21[4];
    for (int k = 0; k < 4; ++k) pf1[k] = project2(f1.v[k], b0, b1);
    for (int k = 0; k < 4; ++k) v21[k] = iso_apply(pos12, f2.v[k]);
    for (int k = 0; k < 4; ++k) pf2[k] = project2(v21[k], b0, b1);
    if (f2.n > 2) {
        V3<T> normal2_1 = cross(v21[2] - v21[1], v21[0] - v21[1]);
        T denom = dot(normal2_1, sep);
        if (!rel_eq(denom, T(0))) {
            const int last2 = f2.n - 1;
            for (int i = 0; i < f1.n; ++i) {
                P2<T> p1 = pf1[i];
                T sign = perp2(sub2(pf2[0], pf2[last2]), sub2(p1, pf2[last2]));
                bool outside = false;
                for (int j = 0; j < last2; ++j) {
                    T ns = perp2(sub2(pf2[j + 1], pf2[j]), sub2(p1, pf2[j]));
                    if (sign == T(0)) {
                        sign = ns;
                    } else if (sign * ns < T(0)) {
                        outside = true;
                        break;
                    }
                }
                if (outside) continue;
                T dist = dot(v21[0] - f1.v[i], normal2_1) / denom;
                V3<T> lp1 = f1.v[i];
                V3<T> lp21 = f1.v[i] + sep * dist;
                out.push_back(TrackedContact<T>{lp1, iso_inv_point(pos12, lp21), f1.vids[i], f2.fid, dist});
            }
        }
    }
    if (f1.n > 2) {
        V3<T> normal1 = cross(f1.v[2] - f1.v[1], f1.v[0] - f1.v[1]);
        T denom = -dot(normal1, sep);
        if (!rel_eq(denom, T(0))) {
            const int last1 = f1.n - 1;
            for (int i = 0; i < f2.n; ++i) {
                P2<T> p2 = pf2[i];
                T sign = perp2(sub2(pf1[0], pf1[last1]), sub2(p2, pf1[last1]));
                bool outside = false;
                for (int j = 0; j < last1; ++j) {
                    T ns = perp2(sub2(pf1[j + 1], pf1[j]), sub2(p2, pf1[j]));
                    if (sign == T(0)) {
                        sign = ns;
                    } else if (sign * ns < T(0)) {
                        outside = true;
                        break;
                    }
                }
                if (outside) continue;
                T dist = dot(f1.v[0] - v21[i], normal1) / denom;
                V3<T> lp21 = v21[i];
                V3<T> lp1 = v21[i] - sep * dist;
                out.push_back(TrackedContact<T>{lp1, iso_inv_point(pos12, lp21), f1.fid, f2.vids[i], dist});
            }
        }
    }
    for (int j = 0; j < f2.n; ++j) {
        P2<T> e20 = pf2[j], e21 = pf2[(j + 1) % f2.n];
        for (int i = 0; i < f1.n; ++i) {
            P2<T> e10 = pf1[i], e11 = pf1[(i + 1) % f1.n];
            T s, t;
            if (closest_points_line2d(e10, e11, e20, e21, s, t)) {
                if (s > T(0) && s < T(1) && t > T(0) && t < T(1)) {
                    V3<T> lp1 = f1.v[i] * (T(1) - s) + f1.v[(i + 1) % f1.n] * s;
                    V3<T> lp21 = v21[j] * (T(1) - t) + v21[(j + 1) % f2.n] * t;
                    T dist = dot(lp21 - lp1, sep);
                    out.push_back(TrackedContact<T>{lp1, iso_inv_point(pos12, lp21), f1.eids[i], f2.eids[j], dist});
                }
            }
        }
    }
}

// ContactManifold::try_update_contacts_eps (contact_manifold.rs:662-699).
template <class T> bool try_update_contacts_eps(Manifold<T> &m, const Iso<T> &pos12, T angle_dot_threshold,
                                                T dist_sq_threshold) {
    if (m.points.empty()) return false;
    const V3<T> local_n2 = rotate(pos12.q, m.n2);
    if (-dot(m.n1, local_n2) < angle_dot_threshold) return false;
    for (auto &pt : m.points) {
        const V3<T> local_p2 = transform_point(pos12, pt.p2);
        const V3<T> dpt = local_p2 - pt.p1;
        const T dist = dot(dpt, m.n1);
        if (dist * pt.dist < T(0)) return false;
        const V3<T> new_p1 = local_p2 - m.n1 * dist;
        const V3<T> d = new_p1 - pt.p1;
        if (dot(d, d) > dist_sq_threshold) return false;
        pt.dist = dist;
        pt.p1 = new_p1;
    }
    return true;
}

template <class T> constexpr T cos_1_degrees() { return T(0.99984769515); }  // utils/consts.rs:9

// contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162) on the manifold
// `m` of the previous frame (no points and a zero normal for a fresh one), with
// no normal constraints.  Returns 2 if try_update_contacts_eps kept it, 1 for
// ClosestPoints (points filled), 0 for NoIntersection (n1 = GJK's separating
// direction, the cached normal), -1 if a shape has no feature map here.  The
// per-point data match_contacts transfers is not modelled.
template <class T>
int manifold_pfm_pfm(const Iso<T> &pos12, const Shape<T> &s1, const Shape<T> &s2, T prediction, Manifold<T> &m) {
    Shape<T> pfm1, pfm2;
    T br1, br2;
    if (!pfm_of(s1, pfm1, br1) || !pfm_of(s2, pfm2, br2)) {
        m.points.clear();
        m.n1 = m.n2 = v3(T(0), T(0), T(0));
        return -1;
    }
    if (try_update_contacts_eps(m, pos12, cos_1_degrees<T>(), T(1.0e-6))) return 2;
    // Unit::try_new(local_n1, DEFAULT_EPSILON)
    const T e = Eps<T>::eps;
    const T sq = dot(m.n1, m.n1);
    V3<T> init{};
    const bool seeded = sq > e * e;
    if (seeded) init = m.n1 / std::sqrt(sq);
    m.points.clear();  // ContactManifold::clear (contact_manifold.rs:793-795) keeps both normals
    const T total_prediction = prediction + br1 + br2;
    Simplex<T> simplex;
    GjkResult<T> r = contact_sm_sm_with_params(pos12, pfm1, pfm2, total_prediction, simplex,
                                               seeded ? &init : static_cast<const V3<T> *>(nullptr), nullptr, false);
    if (r.kind == GJK_CLOSEST_POINTS) {
        const V3<T> p1 = r.p1, p21 = r.p2, dir = r.n;
        V3<T> n1 = dir;
        V3<T> n2 = inv_rotate(pos12.q, -dir);
        PolyFeature<T> f1 = support_feature(pfm1, n1), f2 = support_feature(pfm2, n2);
        if (f1.n == 2 && f2.n == 2) contacts_edge_edge(pos12, f1, n1, f2, m.points);
        else contacts_face_face(pos12, f1, n1, f2, m.points);
        m.points.push_back(TrackedContact<T>{p1, iso_inv_point(pos12, p21), PF_UNKNOWN, PF_UNKNOWN, dot(p21 - p1, n1)});
        if (br1 != T(0) || br2 != T(0)) {
            for (auto &c : m.points) {
                c.p1 = c.p1 + n1 * br1;
                c.p2 = c.p2 + n2 * br2;
                c.dist -= br1 + br2;
            }
        }
        m.n1 = n1;
        m.n2 = n2;
        return 1;
    }
    // NoIntersection caches its direction in local_n1, anything else resets it;
    // local_n2 keeps the previous frame's value (contact_manifolds_pfm_pfm.rs:150-157)
    m.n1 = r.kind == GJK_NO_INTERSECTION ? r.n : v3(T(0), T(0), T(0));
    return 0;
}

}  // namespace oracle
