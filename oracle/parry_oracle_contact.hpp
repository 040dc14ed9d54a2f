// parry_oracle_contact.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of parry3d's narrowphase contact path for support-map shapes:
// support maps, CSO points, the 3D Voronoi simplex, GJK, EPA and
// contact_support_map_support_map, plus query::contact's frame handling.
// Template over T (float = parry3d, double = parry3d-f64); compiled with
// -ffp-contract=off.  Citations are dimforge/parry 0.25.3 path:line.
//
// Third-party behaviour restated from published sources (unpinned at the ulp):
//   nalgebra Unit::try_new_and_get(v, m): |v|^2 > m^2 ? (v/|v|, |v|) : None
//   nalgebra try_normalize(m): |v| <= m ? None : v/|v|
//   nalgebra orthonormal_subspace_basis([v]) first vector: a x v with
//     a = normalize(|v.x| > |v.y| ? (v.z, 0, -v.x) : (0, -v.z, v.y))
//   approx relative_eq!(x, 0, epsilon = e)  <=>  |x| <= e (finite x)
//   std BinaryHeap push/pop (sift_up / sift_down_to_bottom + sift_up)
#pragma once

#include <cstdint>
#include <vector>

#include "parry_oracle.hpp"

namespace oracle {

template <class T> struct Eps;
template <> struct Eps<float> {
    static constexpr float eps = 1.1920929e-07f;  // f32::EPSILON (src/lib.rs:102)
};
template <> struct Eps<double> {
    static constexpr double eps = 2.220446049250313e-16;
};

// TRIANGLE (pts[0..3]) is parry's Triangle (src/shape/triangle.rs): a TriMesh
// part in composite contacts, and the shape of the reference's triangle EPA
// known-answer test (crates/parry3d/tests/geometry/epa3.rs:25-59).
// ORIGIN is parry's ConstantOrigin (special_support_maps.rs:32-54), the second
// operand of GJK/EPA point projections.
// SEGMENT (internal) is a Capsule's segment (0, -h, 0)-(0, h, 0), the feature map
// contact manifolds use for a Capsule (shape.rs:855-857).
enum ShapeType : uint32_t {
    BALL = 0, CUBOID = 1, CAPSULE = 2, CONE = 3, CONVEX_POLYHEDRON = 4, TRIANGLE = 5, ORIGIN = 6, SEGMENT = 7
};

// A ConvexPolyhedron face (convex_polyhedron.rs:59-63): normal and the range of
// its vertices and edges in the face lists.
template <class T> struct PolyFace {
    V3<T> normal;
    uint32_t first, count;
};

template <class T> struct Shape {
    uint32_t type;
    const V3<T> *pts;  // ConvexPolyhedron vertices
    uint32_t n;
    T p[4];  // Ball: r; Cuboid: half extents; Capsule: half_height, radius; Cone: half_height, radius
    // ConvexPolyhedron faces (contact manifolds only): nfaces == 0 means none.
    const PolyFace<T> *faces = nullptr;
    uint32_t nfaces = 0;
    const uint32_t *fverts = nullptr, *fedges = nullptr;
};

template <class T> inline T copy_sign(T from, T to) { return std::copysign(to, from); }

// SupportMap::local_support_point for the shapes on the path.
template <class T> V3<T> local_support_point(const Shape<T> &s, V3<T> dir) {
    switch (s.type) {
        case CONVEX_POLYHEDRON: {  // point_cloud_support_point.rs:5-19: first strict maximum
            uint32_t best = 0;
            T best_dot = dot(s.pts[0], dir);
            for (uint32_t i = 1; i < s.n; ++i) {
                T d = dot(s.pts[i], dir);
                if (d > best_dot) {
                    best_dot = d;
                    best = i;
                }
            }
            return s.pts[best];
        }
        case TRIANGLE: {  // triangle.rs:697-715
            T d1 = dot(s.pts[0], dir), d2 = dot(s.pts[1], dir), d3 = dot(s.pts[2], dir);
            if (d1 > d2) return d1 > d3 ? s.pts[0] : s.pts[2];
            return d2 > d3 ? s.pts[1] : s.pts[2];
        }
        case BALL: {  // ball.rs:242-249: Unit::new_normalize(dir) * radius
            return normalize(dir) * s.p[0];
        }
        case CUBOID: {  // cuboid.rs:454-456, wops.rs:30-35: sign bit of dir copied onto half extents
            return v3(copy_sign(dir.x, s.p[0]), copy_sign(dir.y, s.p[1]), copy_sign(dir.z, s.p[2]));
        }
        case CAPSULE: {  // capsule.rs:407-418, segment a = (0,-h,0), b = (0,h,0)
            V3<T> d;
            T sq = dot(dir, dir);
            if (sq > T(0)) d = dir / std::sqrt(sq);
            else d = v3(T(0), T(1), T(0));
            V3<T> a = v3(T(0), -s.p[0], T(0)), b = v3(T(0), s.p[0], T(0));
            if (dot(d, a) > dot(d, b)) return a + d * s.p[1];
            return b + d * s.p[1];
        }
        case ORIGIN:  // special_support_maps.rs:46-48
            return v3(T(0), T(0), T(0));
        case SEGMENT: {  // segment.rs:461-467
            V3<T> a = v3(T(0), -s.p[0], T(0)), b = v3(T(0), s.p[0], T(0));
            return dot(a, dir) > dot(b, dir) ? a : b;
        }
        case CONE: {  // cone.rs:162-182
            V3<T> v = dir;
            v.y = T(0);
            T nrm = norm(v);
            v = v / nrm;
            if (nrm == T(0)) {
                v = v3(T(0), copy_sign(dir.y, s.p[0]), T(0));
            } else {
                v = v * s.p[1];
                v.y = -s.p[0];
                if (dot(dir, v) < dir.y * s.p[0]) v = v3(T(0), s.p[0], T(0));
            }
            return v;
        }
    }
    return v3(T(0), T(0), T(0));
}

// SupportMap::support_point (support_map.rs:380-383); ConstantOrigin
// overrides it with m's translation (special_support_maps.rs:36-38).
template <class T> V3<T> support_point(const Iso<T> &m, const Shape<T> &s, V3<T> dir) {
    if (s.type == ORIGIN) return m.t;
    V3<T> local_dir = inv_rotate(m.q, dir);
    return transform_point(m, local_support_point(s, local_dir));
}

template <class T> struct CSO {  // cso_point.rs:12-89
    V3<T> point, orig1, orig2;
};
template <class T> CSO<T> cso_from_shapes(const Iso<T> &pos12, const Shape<T> &g1, const Shape<T> &g2, V3<T> dir) {
    V3<T> sp1 = local_support_point(g1, dir);
    V3<T> sp2 = support_point(pos12, g2, -dir);
    return CSO<T>{sp1 - sp2, sp1, sp2};
}

// relative_eq!(p, origin) with approx's defaults (epsilon = max_relative = EPS).
template <class T> inline bool rel_eq_origin(V3<T> p) {
    const T e = Eps<T>::eps;
    return std::fabs(p.x) <= e && std::fabs(p.y) <= e && std::fabs(p.z) <= e;
}

// ---- point projections with location ---------------------------------------

enum LocKind { ON_VERTEX, ON_EDGE, ON_FACE, ON_SOLID };
template <class T> struct Loc {
    LocKind kind;
    uint32_t id;  // vertex / edge / face index
    T bc[3];      // barycentric coordinates (edge: 2, face: 3)
};

// Segment (point_segment.rs:45-85), projecting pt.
template <class T> V3<T> project_segment(V3<T> a, V3<T> b, V3<T> pt, Loc<T> &loc) {
    V3<T> ab = b - a;
    V3<T> ap = pt - a;
    T ab_ap = dot(ab, ap);
    T sqnab = dot(ab, ab);
    if (ab_ap <= T(0)) {
        loc = Loc<T>{ON_VERTEX, 0, {T(1), T(0), T(0)}};
        return a;
    } else if (ab_ap >= sqnab) {
        loc = Loc<T>{ON_VERTEX, 1, {T(1), T(0), T(0)}};
        return b;
    }
    T u = ab_ap / sqnab;
    loc = Loc<T>{ON_EDGE, 0, {T(1) - u, u, T(0)}};
    return a + ab * u;
}

// Triangle (point_triangle.rs:54-291), solid = true.  *inside = relative_eq!(proj, pt).
template <class T> V3<T> project_triangle(V3<T> a, V3<T> b, V3<T> c, V3<T> pt, Loc<T> &loc, bool *inside) {
    V3<T> ab = b - a, ac = c - a, ap = pt - a;
    T ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
    auto res = [&](V3<T> proj) {
        *inside = rel_eq_origin(proj - pt);
        return proj;
    };
    if (ab_ap <= T(0) && ac_ap <= T(0)) {
        loc = Loc<T>{ON_VERTEX, 0, {}};
        return res(a);
    }
    V3<T> bp = pt - b;
    T ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
    if (ab_bp >= T(0) && ac_bp <= ab_bp) {
        loc = Loc<T>{ON_VERTEX, 1, {}};
        return res(b);
    }
    V3<T> cp = pt - c;
    T ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
    if (ac_cp >= T(0) && ab_cp <= ac_cp) {
        loc = Loc<T>{ON_VERTEX, 2, {}};
        return res(c);
    }
    V3<T> bc = c - b;
    // stable_check_edges_voronoi (dim3)
    V3<T> n = cross(ab, ac);
    T vc = dot(n, cross(ab, ap));
    if (vc < T(0) && ab_ap >= T(0) && ab_bp <= T(0)) {
        T v = ab_ap / dot(ab, ab);
        loc = Loc<T>{ON_EDGE, 0, {T(1) - v, v, T(0)}};
        return res(a + ab * v);
    }
    T vb = -dot(n, cross(ac, cp));
    if (vb < T(0) && ac_ap >= T(0) && ac_cp <= T(0)) {
        T w = ac_ap / dot(ac, ac);
        loc = Loc<T>{ON_EDGE, 2, {T(1) - w, w, T(0)}};
        return res(a + ac * w);
    }
    T va = dot(n, cross(bc, bp));
    if (va < T(0) && ac_bp - ab_bp >= T(0) && ab_cp - ac_cp >= T(0)) {
        T w = dot(bc, bp) / dot(bc, bc);
        loc = Loc<T>{ON_EDGE, 1, {T(1) - w, w, T(0)}};
        return res(b + bc * w);
    }
    uint32_t clockwise = dot(n, ap) >= T(0) ? 0 : 1;
    if (va + vb + vc != T(0)) {
        T denom = T(1) / (va + vb + vc);
        T v = vb * denom;
        T w = vc * denom;
        loc = Loc<T>{ON_FACE, clockwise, {T(1) - v - w, v, w}};
        return res((a + ab * v) + ac * w);
    }
    loc = Loc<T>{ON_SOLID, 0, {}};
    *inside = true;
    return pt;
}

// TrianglePointLocation::barycentric_coordinates (triangle.rs:121-148).
template <class T> bool tri_bcoords(const Loc<T> &l, T out[3]) {
    out[0] = out[1] = out[2] = T(0);
    switch (l.kind) {
        case ON_VERTEX: out[l.id] = T(1); return true;
        case ON_EDGE: {
            int i0 = l.id == 1 ? 1 : 0, i1 = l.id == 0 ? 1 : 2;
            out[i0] = l.bc[0];
            out[i1] = l.bc[1];
            return true;
        }
        case ON_FACE: out[0] = l.bc[0]; out[1] = l.bc[1]; out[2] = l.bc[2]; return true;
        default: return false;
    }
}

// Tetrahedron (point_tetrahedron.rs:28-340), solid = true.
template <class T> V3<T> project_tetrahedron(V3<T> a, V3<T> b, V3<T> c, V3<T> d, V3<T> pt, Loc<T> &loc) {
    V3<T> ab = b - a, ac = c - a, ad = d - a, ap = pt - a;
    T ap_ab = dot(ap, ab), ap_ac = dot(ap, ac), ap_ad = dot(ap, ad);
    if (ap_ab <= T(0) && ap_ac <= T(0) && ap_ad <= T(0)) { loc = Loc<T>{ON_VERTEX, 0, {}}; return a; }
    V3<T> bc = c - b, bd = d - b, bp = pt - b;
    T bp_bc = dot(bp, bc), bp_bd = dot(bp, bd), bp_ab = dot(bp, ab);
    if (bp_bc <= T(0) && bp_bd <= T(0) && bp_ab >= T(0)) { loc = Loc<T>{ON_VERTEX, 1, {}}; return b; }
    V3<T> cd = d - c, cp = pt - c;
    T cp_ac = dot(cp, ac), cp_bc = dot(cp, bc), cp_cd = dot(cp, cd);
    if (cp_cd <= T(0) && cp_bc >= T(0) && cp_ac >= T(0)) { loc = Loc<T>{ON_VERTEX, 2, {}}; return c; }
    V3<T> dp = pt - d;
    T dp_cd = dot(dp, cd), dp_bd = dot(dp, bd), dp_ad = dot(dp, ad);
    if (dp_ad >= T(0) && dp_bd >= T(0) && dp_cd >= T(0)) { loc = Loc<T>{ON_VERTEX, 3, {}}; return d; }

    // check_edge: returns true (and the projection) if pt is in edge i's region.
    auto check_edge = [&](uint32_t i, V3<T> ea, V3<T> nabc, V3<T> nabd, V3<T> eap, V3<T> eab, T eap_ab, T ebp_ab,
                          T &dabc, T &dabd, V3<T> &out) -> bool {
        T ab_ab = eap_ab - ebp_ab;
        V3<T> ap_x_ab = cross(eap, eab);
        dabc = dot(ap_x_ab, nabc);
        dabd = dot(ap_x_ab, nabd);
        if (ab_ab != T(0) && dabc >= T(0) && dabd >= T(0) && eap_ab >= T(0) && eap_ab <= ab_ab) {
            T u = eap_ab / ab_ab;
            loc = Loc<T>{ON_EDGE, i, {T(1) - u, u, T(0)}};
            out = ea + eab * u;
            return true;
        }
        return false;
    };
    V3<T> out;
    V3<T> nabc = cross(ab, ac), nabd = cross(ab, ad);
    T dabc, dabd, dacd, dacb, dadb, dadc, dbca, dbcd, dbdc, dbda, dcda, dcdb;
    if (check_edge(0, a, nabc, nabd, ap, ab, ap_ab, bp_ab, dabc, dabd, out)) return out;
    V3<T> nacd = cross(ac, ad);
    if (check_edge(1, a, nacd, -nabc, ap, ac, ap_ac, cp_ac, dacd, dacb, out)) return out;
    if (check_edge(2, a, -nabd, -nacd, ap, ad, ap_ad, dp_ad, dadb, dadc, out)) return out;
    V3<T> nbcd = cross(bc, bd);
    if (check_edge(3, b, nabc, nbcd, bp, bc, bp_bc, cp_bc, dbca, dbcd, out)) return out;
    if (check_edge(4, b, -nbcd, nabd, bp, bd, bp_bd, dp_bd, dbdc, dbda, out)) return out;
    if (check_edge(5, c, nacd, nbcd, cp, cd, cp_cd, dp_cd, dcda, dcdb, out)) return out;

    auto check_face = [&](uint32_t i, V3<T> fa, V3<T> fb, V3<T> fc, V3<T> fap, V3<T> fbp, V3<T> fcp, V3<T> fab,
                          V3<T> fac, V3<T> fad, T d1, T d2, T d3, V3<T> &res) -> bool {
        if (d1 < T(0) && d2 < T(0) && d3 < T(0)) {
            V3<T> n = cross(fab, fac);
            if (dot(n, fad) * dot(n, fap) < T(0)) {
                T nn = norm(n);
                if (nn <= Eps<T>::eps) return false;  // try_normalize(DEFAULT_EPSILON) failed
                V3<T> normal = n / nn;
                T vc = dot(normal, cross(fap, fbp));
                T va = dot(normal, cross(fbp, fcp));
                T vb = dot(normal, cross(fcp, fap));
                T denom = va + vb + vc;
                T inv = T(1) / denom;
                T bc0 = va * inv, bc1 = vb * inv, bc2 = vc * inv;
                loc = Loc<T>{ON_FACE, i, {bc0, bc1, bc2}};
                res = (fa * bc0 + fb * bc1) + fc * bc2;
                return true;
            }
        }
        return false;
    };
    if (check_face(0, a, b, c, ap, bp, cp, ab, ac, ad, dabc, dbca, dacb, out)) return out;
    if (check_face(1, a, b, d, ap, bp, dp, ab, ad, ac, dadb, dabd, dbda, out)) return out;
    if (check_face(2, a, c, d, ap, cp, dp, ac, ad, ab, dacd, dcda, dadc, out)) return out;
    if (check_face(3, b, c, d, bp, cp, dp, bc, bd, -ab, dbcd, dcdb, dbdc, out)) return out;
    loc = Loc<T>{ON_SOLID, 0, {}};
    return pt;
}

// ---- VoronoiSimplex (voronoi_simplex3.rs:13-351) ----------------------------

template <class T> struct Simplex {
    size_t prev_vertices[4] = {0, 1, 2, 3};
    T prev_proj[3] = {0, 0, 0};
    size_t prev_dim = 0;
    CSO<T> vertices[4] = {};
    T proj[3] = {0, 0, 0};
    size_t dim = 0;

    void swap(size_t i1, size_t i2) {
        std::swap(vertices[i1], vertices[i2]);
        std::swap(prev_vertices[i1], prev_vertices[i2]);
    }
    void reset(const CSO<T> &pt) {
        dim = 0;
        prev_dim = 0;
        vertices[0] = pt;
    }
    bool add_point(const CSO<T> &pt) {
        const T eps_tol = Eps<T>::eps * T(10);
        prev_dim = dim;
        for (int i = 0; i < 3; ++i) prev_proj[i] = proj[i];
        for (size_t i = 0; i < 4; ++i) prev_vertices[i] = i;
        if (dim == 0) {
            V3<T> d = vertices[0].point - pt.point;
            if (dot(d, d) < eps_tol) return false;
        } else if (dim == 1) {
            V3<T> ab = vertices[1].point - vertices[0].point;
            V3<T> ac = pt.point - vertices[0].point;
            V3<T> c = cross(ab, ac);
            if (dot(c, c) < eps_tol) return false;
        } else {  // dim == 2
            V3<T> ab = vertices[1].point - vertices[0].point;
            V3<T> ac = vertices[2].point - vertices[0].point;
            V3<T> ap = pt.point - vertices[0].point;
            V3<T> n = normalize(cross(ab, ac));
            if (std::fabs(dot(n, ap)) < eps_tol) return false;
        }
        dim += 1;
        vertices[dim] = pt;
        return true;
    }

    V3<T> project_origin_and_reduce() {
        const V3<T> o = v3(T(0), T(0), T(0));
        if (dim == 0) {
            proj[0] = T(1);
            return vertices[0].point;
        } else if (dim == 1) {
            Loc<T> l;
            V3<T> p = project_segment(vertices[0].point, vertices[1].point, o, l);
            if (l.kind == ON_VERTEX) {
                if (l.id == 1) swap(0, 1);
                proj[0] = T(1);
                dim = 0;
            } else {
                proj[0] = l.bc[0];
                proj[1] = l.bc[1];
            }
            return p;
        } else if (dim == 2) {
            Loc<T> l;
            bool inside;
            V3<T> p = project_triangle(vertices[0].point, vertices[1].point, vertices[2].point, o, l, &inside);
            if (l.kind == ON_VERTEX) {
                swap(0, l.id);
                proj[0] = T(1);
                dim = 0;
            } else if (l.kind == ON_EDGE) {
                if (l.id == 0) {
                    proj[0] = l.bc[0];
                    proj[1] = l.bc[1];
                } else if (l.id == 1) {
                    swap(0, 2);
                    proj[0] = l.bc[1];
                    proj[1] = l.bc[0];
                } else {
                    swap(1, 2);
                    proj[0] = l.bc[0];
                    proj[1] = l.bc[1];
                }
                dim = 1;
            } else if (l.kind == ON_FACE) {
                proj[0] = l.bc[0];
                proj[1] = l.bc[1];
                proj[2] = l.bc[2];
            }
            return p;
        } else {
            Loc<T> l;
            V3<T> p = project_tetrahedron(vertices[0].point, vertices[1].point, vertices[2].point, vertices[3].point,
                                          o, l);
            if (l.kind == ON_VERTEX) {
                swap(0, l.id);
                proj[0] = T(1);
                dim = 0;
            } else if (l.kind == ON_EDGE) {
                switch (l.id) {
                    case 0: break;
                    case 1: swap(1, 2); break;
                    case 2: swap(1, 3); break;
                    case 3: swap(0, 2); break;
                    case 4: swap(0, 3); break;
                    case 5: swap(0, 2); swap(1, 3); break;
                }
                if (l.id == 3 || l.id == 4) {
                    proj[0] = l.bc[1];
                    proj[1] = l.bc[0];
                } else {
                    proj[0] = l.bc[0];
                    proj[1] = l.bc[1];
                }
                dim = 1;
            } else if (l.kind == ON_FACE) {
                switch (l.id) {
                    case 0:
                        proj[0] = l.bc[0]; proj[1] = l.bc[1]; proj[2] = l.bc[2];
                        break;
                    case 1:
                        vertices[2] = vertices[3];
                        proj[0] = l.bc[0]; proj[1] = l.bc[1]; proj[2] = l.bc[2];
                        break;
                    case 2:
                        vertices[1] = vertices[3];
                        proj[0] = l.bc[0]; proj[1] = l.bc[2]; proj[2] = l.bc[1];
                        break;
                    case 3:
                        vertices[0] = vertices[3];
                        proj[0] = l.bc[2]; proj[1] = l.bc[0]; proj[2] = l.bc[1];
                        break;
                }
                dim = 2;
            }
            return p;
        }
    }
};

// ---- GJK (gjk.rs:353-453, 797-818) -----------------------------------------

enum GjkKind { GJK_INTERSECTION, GJK_CLOSEST_POINTS, GJK_NO_INTERSECTION };
template <class T> struct GjkResult {
    GjkKind kind;
    V3<T> p1, p2, n;
};

template <class T> void gjk_result(const Simplex<T> &s, bool prev, V3<T> &p1, V3<T> &p2) {
    p1 = p2 = v3(T(0), T(0), T(0));
    if (prev) {
        for (size_t i = 0; i < s.prev_dim + 1; ++i) {
            T c = s.prev_proj[i];
            const CSO<T> &pt = s.vertices[s.prev_vertices[i]];
            p1 = p1 + pt.orig1 * c;
            p2 = p2 + pt.orig2 * c;
        }
    } else {
        for (size_t i = 0; i < s.dim + 1; ++i) {
            T c = s.proj[i];
            const CSO<T> &pt = s.vertices[i];
            p1 = p1 + pt.orig1 * c;
            p2 = p2 + pt.orig2 * c;
        }
    }
}

template <class T>
GjkResult<T> gjk_closest_points(const Iso<T> &pos12, const Shape<T> &g1, const Shape<T> &g2, T max_dist,
                                Simplex<T> &simplex, uint32_t *iters) {
    const T eps_tol = Eps<T>::eps * T(10);
    const T eps_rel = std::sqrt(eps_tol);
    V3<T> proj = simplex.project_origin_and_reduce();
    V3<T> old_dir;
    {
        T sq = dot(proj, proj);
        if (!(sq > T(0))) return GjkResult<T>{GJK_INTERSECTION, {}, {}, {}};
        old_dir = -(proj / std::sqrt(sq));
    }
    T max_bound = Consts<T>::max_value;
    V3<T> dir;
    uint32_t niter = 0;
    if (iters) *iters = 0;
    for (;;) {
        T old_max_bound = max_bound;
        V3<T> mp = -proj;
        T sq = dot(mp, mp);
        if (sq > eps_tol * eps_tol) {
            T nrm = std::sqrt(sq);
            dir = mp / nrm;
            max_bound = nrm;
        } else {
            return GjkResult<T>{GJK_INTERSECTION, {}, {}, {}};
        }
        if (max_bound >= old_max_bound) {
            GjkResult<T> r{GJK_CLOSEST_POINTS, {}, {}, old_dir};
            gjk_result(simplex, true, r.p1, r.p2);
            return r;
        }
        CSO<T> cso = cso_from_shapes(pos12, g1, g2, dir);
        T min_bound = -dot(dir, cso.point);
        if (!std::isfinite(min_bound)) return GjkResult<T>{GJK_NO_INTERSECTION, {}, {}, v3(T(1), T(0), T(0))};
        if (min_bound > max_dist) {
            return GjkResult<T>{GJK_NO_INTERSECTION, {}, {}, dir};
        } else if (max_bound - min_bound <= eps_rel * max_bound) {
            GjkResult<T> r{GJK_CLOSEST_POINTS, {}, {}, dir};
            gjk_result(simplex, false, r.p1, r.p2);
            return r;
        }
        if (!simplex.add_point(cso)) {
            GjkResult<T> r{GJK_CLOSEST_POINTS, {}, {}, dir};
            gjk_result(simplex, false, r.p1, r.p2);
            return r;
        }
        old_dir = dir;
        proj = simplex.project_origin_and_reduce();
        if (simplex.dim == 3) {
            if (min_bound >= eps_tol) {
                GjkResult<T> r{GJK_CLOSEST_POINTS, {}, {}, old_dir};
                gjk_result(simplex, true, r.p1, r.p2);
                return r;
            }
            return GjkResult<T>{GJK_INTERSECTION, {}, {}, {}};
        }
        niter += 1;
        if (iters) *iters = niter;
        if (niter == 100) return GjkResult<T>{GJK_NO_INTERSECTION, {}, {}, v3(T(1), T(0), T(0))};
    }
}

// ---- EPA (epa3.rs:17-675) ----------------------------------------------------

template <class T> struct EpaFaceId {
    size_t id;
    T neg_dist;
};
// Ord of FaceId (epa3.rs:40-53): by neg_dist; NaN compares Equal.
template <class T> inline int face_cmp(const EpaFaceId<T> &a, const EpaFaceId<T> &b) {
    if (a.neg_dist < b.neg_dist) return -1;
    if (a.neg_dist > b.neg_dist) return 1;
    return 0;
}

// std::collections::BinaryHeap (max-heap) with the standard library's sift order.
template <class T> struct Heap {
    std::vector<EpaFaceId<T>> d;
    size_t sift_up(size_t start, size_t pos) {
        EpaFaceId<T> e = d[pos];
        while (pos > start) {
            size_t parent = (pos - 1) / 2;
            if (face_cmp(e, d[parent]) <= 0) break;
            d[pos] = d[parent];
            pos = parent;
        }
        d[pos] = e;
        return pos;
    }
    void push(const EpaFaceId<T> &x) {
        d.push_back(x);
        sift_up(0, d.size() - 1);
    }
    void sift_down_to_bottom(size_t pos) {
        size_t end = d.size();
        size_t start = pos;
        EpaFaceId<T> e = d[pos];
        size_t child = 2 * pos + 1;
        while (child + 2 <= end) {  // child <= end - 2
            if (face_cmp(d[child], d[child + 1]) <= 0) child += 1;
            d[pos] = d[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (child == end - 1) {
            d[pos] = d[child];
            pos = child;
        }
        d[pos] = e;
        sift_up(start, pos);
    }
    bool pop(EpaFaceId<T> &out) {
        if (d.empty()) return false;
        EpaFaceId<T> item = d.back();
        d.pop_back();
        if (!d.empty()) {
            std::swap(item, d[0]);
            sift_down_to_bottom(0);
        }
        out = item;
        return true;
    }
};

template <class T> struct EpaFace {
    size_t pts[3];
    size_t adj[3];
    V3<T> normal;
    T bcoords[3];
    bool deleted;
};

// How EPA::closest_points ended (recorded for tests and error reports only).
// EPA_STUCK is the second clause of epa3.rs:568-572.  It needs
// candidate_max_dist < max_dist right after max_dist = min(max_dist,
// candidate_max_dist) (epa3.rs:560-563), so it cannot hold: it is counted
// separately to show that no pair ever takes it.
enum EpaExit : uint32_t {
    EPA_NOT_RUN = 0,
    EPA_CONVERGED = 1,   // max_dist - curr_dist < 100 EPS (epa3.rs:568)
    EPA_STUCK = 2,       // |curr_dist - old_dist| < EPS && candidate < max_dist (epa3.rs:569-572)
    EPA_NUMERICAL = 3,   // a new face closer than the popped one (epa3.rs:614-621): returns the popped face
    EPA_ITER_CAP = 4,    // more than 100 iterations (epa3.rs:640-645): returns the best face
    EPA_HEAP_EMPTY = 5,  // every face popped (epa3.rs:547, 648-650): returns the best face
    EPA_FAILED = 6,      // None (no face, silhouette empty, all silhouette faces deleted, FaceId::new)
    EPA_POINT = 7,       // simplex of dimension 0 (epa3.rs:449-453)
};

template <class T> struct Epa {
    uint32_t exit = EPA_NOT_RUN;
    std::vector<CSO<T>> vertices;
    std::vector<EpaFace<T>> faces;
    std::vector<std::pair<size_t, size_t>> silhouette;  // (face_id, opp_pt_id)
    Heap<T> heap;

    static bool ccw_face_normal(V3<T> a, V3<T> b, V3<T> c, V3<T> &n) {  // ccw_face_normal.rs:19-27
        V3<T> r = cross(b - a, c - a);
        T sq = dot(r, r);
        const T e = Eps<T>::eps;
        if (sq > e * e) {
            n = r / std::sqrt(sq);
            return true;
        }
        return false;
    }
    EpaFace<T> face_with_proj(const T bc[3], const size_t pts[3], const size_t adj[3]) {
        EpaFace<T> f;
        for (int i = 0; i < 3; ++i) {
            f.pts[i] = pts[i];
            f.adj[i] = adj[i];
            f.bcoords[i] = bc[i];
        }
        if (!ccw_face_normal(vertices[pts[0]].point, vertices[pts[1]].point, vertices[pts[2]].point, f.normal))
            f.normal = v3(T(0), T(0), T(0));
        f.deleted = false;
        return f;
    }
    // Face::new (epa3.rs:96-118): (face, proj_inside).
    std::pair<EpaFace<T>, bool> face_new(const size_t pts[3], const size_t adj[3]) {
        Loc<T> l;
        bool inside;
        V3<T> p = project_triangle(vertices[pts[0]].point, vertices[pts[1]].point, vertices[pts[2]].point,
                                   v3(T(0), T(0), T(0)), l, &inside);
        if (l.kind == ON_VERTEX || l.kind == ON_EDGE) {
            T bc[3];
            tri_bcoords(l, bc);
            const T eps_tol = Eps<T>::eps * T(100);
            bool in = inside || dot(p, p) < eps_tol * eps_tol;
            return {face_with_proj(bc, pts, adj), in};
        } else if (l.kind == ON_FACE) {
            return {face_with_proj(l.bc, pts, adj), true};
        }
        const T zero[3] = {T(0), T(0), T(0)};
        return {face_with_proj(zero, pts, adj), false};
    }
    void closest_points_of(const EpaFace<T> &f, V3<T> &p1, V3<T> &p2) const {
        const CSO<T> &a = vertices[f.pts[0]], &b = vertices[f.pts[1]], &c = vertices[f.pts[2]];
        p1 = (a.orig1 * f.bcoords[0] + b.orig1 * f.bcoords[1]) + c.orig1 * f.bcoords[2];
        p2 = (a.orig2 * f.bcoords[0] + b.orig2 * f.bcoords[1]) + c.orig2 * f.bcoords[2];
    }
    static size_t next_ccw_pt_id(const EpaFace<T> &f, size_t id) {
        if (f.pts[0] == id) return 1;
        if (f.pts[1] == id) return 2;
        return 0;
    }
    bool can_be_seen_by(const EpaFace<T> &f, size_t point, size_t opp) const {
        const T eps_tol = Eps<T>::eps * T(10);
        V3<T> p0 = vertices[f.pts[opp]].point;
        V3<T> p1 = vertices[f.pts[(opp + 1) % 3]].point;
        V3<T> p2 = vertices[f.pts[(opp + 2) % 3]].point;
        V3<T> pt = vertices[point].point;
        if (dot(pt - p0, f.normal) >= -eps_tol) return true;
        // Triangle::is_affinely_dependent (triangle.rs:534-546)
        const T E = Eps<T>::eps * T(100);
        V3<T> c = cross(p2 - p1, pt - p1);
        return dot(c, c) <= E * E;
    }
    void compute_silhouette(size_t point, size_t id, size_t opp) {
        if (faces[id].deleted) return;
        if (!can_be_seen_by(faces[id], point, opp)) {
            silhouette.push_back({id, opp});
        } else {
            faces[id].deleted = true;
            size_t a1 = (opp + 2) % 3, a2 = opp;
            size_t adj1 = faces[id].adj[a1], adj2 = faces[id].adj[a2];
            size_t o1 = next_ccw_pt_id(faces[adj1], faces[id].pts[a1]);
            size_t o2 = next_ccw_pt_id(faces[adj2], faces[id].pts[a2]);
            compute_silhouette(point, adj1, o1);
            compute_silhouette(point, adj2, o2);
        }
    }

    // EPA::closest_points (epa3.rs:428-651).
    bool closest_points(const Iso<T> &pos12, const Shape<T> &g1, const Shape<T> &g2, const Simplex<T> &simplex,
                        V3<T> &p1, V3<T> &p2, V3<T> &n, uint32_t *iters) {
        const T eps = Eps<T>::eps;
        const T eps_tol = eps * T(100);
        vertices.clear();
        faces.clear();
        heap.d.clear();
        silhouette.clear();
        exit = EPA_FAILED;
        if (iters) *iters = 0;
        for (size_t i = 0; i < simplex.dim + 1; ++i) vertices.push_back(simplex.vertices[i]);
        auto fid_new = [&](size_t id, T neg_dist, EpaFaceId<T> &out) -> bool {
            if (neg_dist > Eps<T>::eps * T(10)) return false;
            out = EpaFaceId<T>{id, neg_dist};
            return true;
        };
        if (simplex.dim == 0) {
            p1 = p2 = v3(T(0), T(0), T(0));
            n = v3(T(0), T(1), T(0));
            exit = EPA_POINT;
            return true;
        } else if (simplex.dim == 3) {
            V3<T> dp1 = vertices[1].point - vertices[0].point;
            V3<T> dp2 = vertices[2].point - vertices[0].point;
            V3<T> dp3 = vertices[3].point - vertices[0].point;
            if (dot(cross(dp1, dp2), dp3) > T(0)) std::swap(vertices[1], vertices[2]);
            const size_t P[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {0, 3, 1}};
            const size_t A[4][3] = {{3, 1, 2}, {3, 2, 0}, {0, 1, 3}, {2, 1, 0}};
            bool inside[4];
            for (int k = 0; k < 4; ++k) {
                auto fr = face_new(P[k], A[k]);
                faces.push_back(fr.first);
                inside[k] = fr.second;
            }
            for (int k = 0; k < 4; ++k) {
                if (inside[k]) {
                    T dist = dot(faces[k].normal, vertices[k].point);
                    EpaFaceId<T> f;
                    if (!fid_new((size_t)k, -dist, f)) return false;
                    heap.push(f);
                }
            }
            if (!(inside[0] || inside[1] || inside[2] || inside[3])) return false;
        } else {
            if (simplex.dim == 1) {
                V3<T> dpt = vertices[1].point - vertices[0].point;
                V3<T> a = std::fabs(dpt.x) > std::fabs(dpt.y) ? v3(dpt.z, T(0), -dpt.x) : v3(T(0), -dpt.z, dpt.y);
                a = a / norm(a);
                V3<T> dir = cross(a, dpt);
                vertices.push_back(cso_from_shapes(pos12, g1, g2, dir));
            }
            const size_t p1i[3] = {0, 1, 2}, p2i[3] = {0, 2, 1};
            const size_t a1[3] = {1, 1, 1}, a2[3] = {0, 0, 0};
            faces.push_back(face_new(p1i, a1).first);
            faces.push_back(face_new(p2i, a2).first);
            EpaFaceId<T> f;
            if (!fid_new(0, T(0), f)) return false;
            heap.push(f);
            if (!fid_new(1, T(0), f)) return false;
            heap.push(f);
        }
        if (heap.d.empty()) return false;
        uint32_t niter = 0;
        T max_dist = Consts<T>::max_value;
        EpaFaceId<T> best_face_id = heap.d[0];
        T old_dist = T(0);
        EpaFaceId<T> face_id;
        while (heap.pop(face_id)) {
            EpaFace<T> face = faces[face_id.id];
            if (face.deleted) continue;
            CSO<T> cso = cso_from_shapes(pos12, g1, g2, face.normal);
            size_t support_point_id = vertices.size();
            vertices.push_back(cso);
            T candidate_max_dist = dot(cso.point, face.normal);
            if (candidate_max_dist < max_dist) {
                best_face_id = face_id;
                max_dist = candidate_max_dist;
            }
            T curr_dist = -face_id.neg_dist;
            if (max_dist - curr_dist < eps_tol ||
                (std::fabs(curr_dist - old_dist) < eps && candidate_max_dist < max_dist)) {
                const EpaFace<T> &bf = faces[best_face_id.id];
                closest_points_of(bf, p1, p2);
                n = bf.normal;
                exit = max_dist - curr_dist < eps_tol ? EPA_CONVERGED : EPA_STUCK;
                return true;
            }
            old_dist = curr_dist;
            faces[face_id.id].deleted = true;
            size_t o1 = next_ccw_pt_id(faces[face.adj[0]], face.pts[0]);
            size_t o2 = next_ccw_pt_id(faces[face.adj[1]], face.pts[1]);
            size_t o3 = next_ccw_pt_id(faces[face.adj[2]], face.pts[2]);
            compute_silhouette(support_point_id, face.adj[0], o1);
            compute_silhouette(support_point_id, face.adj[1], o2);
            compute_silhouette(support_point_id, face.adj[2], o3);
            size_t first_new_face_id = faces.size();
            if (silhouette.empty()) return false;
            for (const auto &edge : silhouette) {
                if (!faces[edge.first].deleted) {
                    size_t new_face_id = faces.size();
                    size_t opp = edge.second;
                    size_t pt_id1 = faces[edge.first].pts[(opp + 2) % 3];
                    size_t pt_id2 = faces[edge.first].pts[(opp + 1) % 3];
                    const size_t pts[3] = {pt_id1, pt_id2, support_point_id};
                    const size_t adj[3] = {edge.first, new_face_id + 1, new_face_id - 1};
                    auto nf = face_new(pts, adj);
                    faces[edge.first].adj[(opp + 1) % 3] = new_face_id;
                    faces.push_back(nf.first);
                    if (nf.second) {
                        V3<T> pt = vertices[faces[new_face_id].pts[0]].point;
                        T dist = dot(faces[new_face_id].normal, pt);
                        if (dist < curr_dist) {
                            closest_points_of(face, p1, p2);
                            n = face.normal;
                            exit = EPA_NUMERICAL;
                            return true;
                        }
                        EpaFaceId<T> f;
                        if (!fid_new(new_face_id, -dist, f)) return false;
                        heap.push(f);
                    }
                }
            }
            if (first_new_face_id == faces.size()) return false;
            faces[first_new_face_id].adj[2] = faces.size() - 1;
            faces.back().adj[1] = first_new_face_id;
            silhouette.clear();
            niter += 1;
            if (iters) *iters = niter;
            if (niter > 100) {
                exit = EPA_ITER_CAP;
                break;
            }
        }
        if (exit != EPA_ITER_CAP) exit = EPA_HEAP_EMPTY;
        const EpaFace<T> &bf = faces[best_face_id.id];
        closest_points_of(bf, p1, p2);
        n = bf.normal;
        return true;
    }
};

// ---- Shape::compute_aabb for every type on the path ---------------------------
// nalgebra UnitQuaternion::to_rotation_matrix (restated, unpinned at the ulp):
// rows (ww+ii-jj-kk, ij-wk, wj+ik), (wk+ij, ww-ii+jj-kk, jk-wi),
// (ik-wj, wi+jk, ww-ii-jj+kk) with ij = i*j*2 etc.; Matrix3 * v accumulates
// (r0*v0 + r1*v1) + r2*v2.
template <class T> Aabb<T> shape_aabb(const Iso<T> &m, const Shape<T> &s) {
    switch (s.type) {
        case CONVEX_POLYHEDRON:
            return point_cloud_aabb(m, s.pts, s.n);
        case BALL: {  // aabb_ball.rs:7-27
            T r = s.p[0];
            return Aabb<T>{m.t + v3(-r, -r, -r), m.t + v3(r, r, r)};
        }
        case CUBOID: {  // aabb_cuboid.rs:9-14, isometry_ops.rs:14-19
            T i = m.q.x, j = m.q.y, k = m.q.z, w = m.q.w;
            T ww = w * w, ii = i * i, jj = j * j, kk = k * k;
            T ij = i * j * T(2), wk = w * k * T(2), wj = w * j * T(2), ik = i * k * T(2), jk = j * k * T(2),
              wi = w * i * T(2);
            T R[3][3] = {{ww + ii - jj - kk, ij - wk, wj + ik},
                         {wk + ij, ww - ii + jj - kk, jk - wi},
                         {ik - wj, wi + jk, ww - ii - jj + kk}};
            V3<T> he = v3(s.p[0], s.p[1], s.p[2]), ws;
            for (int a = 0; a < 3; ++a)
                ws[a] = (std::fabs(R[a][0]) * he.x + std::fabs(R[a][1]) * he.y) + std::fabs(R[a][2]) * he.z;
            return Aabb<T>{m.t - ws, m.t + ws};  // Aabb::from_half_extents
        }
        case TRIANGLE: {  // Triangle::aabb (aabb_triangle.rs:10-30): transformed vertices, a.min(b).min(c)
            V3<T> a = transform_point(m, s.pts[0]), b = transform_point(m, s.pts[1]), c = transform_point(m, s.pts[2]);
            return Aabb<T>{vmin(vmin(a, b), c), vmax(vmax(a, b), c)};
        }
        case CAPSULE: {  // aabb_capsule.rs:8-20
            V3<T> a = transform_point(m, v3(T(0), -s.p[0], T(0))), b = transform_point(m, v3(T(0), s.p[0], T(0)));
            T r = s.p[1];
            return Aabb<T>{vmin(a, b) - v3(r, r, r), vmax(a, b) + v3(r, r, r)};
        }
        default: {  // support_map_aabb (aabb_utils.rs:10-31)
            Aabb<T> out;
            for (int d = 0; d < 3; ++d) {
                V3<T> basis = v3(T(0), T(0), T(0));
                basis[d] = T(1);
                out.maxs[d] = support_point(m, s, basis)[d];
                basis[d] = T(-1);
                out.mins[d] = support_point(m, s, basis)[d];
            }
            return out;
        }
    }
}

// ---- contact (contact_support_map_support_map.rs, contact_ball_ball.rs,
//      contact_shape_shape.rs) ---------------------------------------------

template <class T> struct Contact {
    V3<T> point1, point2, normal1, normal2;
    T dist;
};

struct ContactStats {
    uint32_t gjk_iters = 0;
    uint32_t epa_iters = 0;
    bool used_epa = false;
    uint32_t epa_exit = EPA_NOT_RUN;  // EpaExit
};

// contact_support_map_support_map_with_params (contact_support_map_support_map.rs:40-77).
// init_dir: nullptr for parry's default seed.
template <class T>
GjkResult<T> contact_sm_sm_with_params(const Iso<T> &pos12, const Shape<T> &g1, const Shape<T> &g2, T prediction,
                                       Simplex<T> &simplex, const V3<T> *init_dir, ContactStats *st,
                                       bool retry_unseeded = true) {
    const T e = Eps<T>::eps;
    V3<T> dflt;
    {
        V3<T> t = pos12.t;
        T sq = dot(t, t);
        if (sq > e * e) dflt = t / std::sqrt(sq);
        else dflt = v3(T(1), T(0), T(0));
    }
    // A given seed replaces the default; a zero (or NaN) seed means the default, and
    // a seeded run that reaches the 100-iteration cap is repeated unseeded (the
    // warm-start contract of this build; parry takes any Unit direction).
    bool seeded = init_dir && dot(*init_dir, *init_dir) > e * e;
    V3<T> dir = seeded ? *init_dir : dflt;
    simplex.reset(cso_from_shapes(pos12, g1, g2, dir));
    uint32_t gi = 0;
    GjkResult<T> c = gjk_closest_points(pos12, g1, g2, prediction, simplex, &gi);
    if (seeded && retry_unseeded && gi == 100) {
        uint32_t gi2 = 0;
        simplex = Simplex<T>();
        simplex.reset(cso_from_shapes(pos12, g1, g2, dflt));
        c = gjk_closest_points(pos12, g1, g2, prediction, simplex, &gi2);
        gi += gi2;
    }
    if (st) st->gjk_iters = gi;
    if (c.kind != GJK_INTERSECTION) return c;
    Epa<T> epa;
    V3<T> p1, p2, n;
    uint32_t ei = 0;
    if (st) st->used_epa = true;
    bool ok = epa.closest_points(pos12, g1, g2, simplex, p1, p2, n, &ei);
    if (st) {
        st->epa_iters = ei;
        st->epa_exit = epa.exit;
    }
    if (ok) return GjkResult<T>{GJK_CLOSEST_POINTS, p1, p2, n};
    return GjkResult<T>{GJK_NO_INTERSECTION, {}, {}, v3(T(1), T(0), T(0))};
}

// contact_support_map_support_map (contact_support_map_support_map.rs:10-32), local frame of shape 1.
template <class T>
bool contact_sm_sm(const Iso<T> &pos12, const Shape<T> &g1, const Shape<T> &g2, T prediction, Contact<T> &out,
                   const V3<T> *init_dir, ContactStats *st) {
    Simplex<T> simplex;
    GjkResult<T> r = contact_sm_sm_with_params(pos12, g1, g2, prediction, simplex, init_dir, st);
    if (r.kind != GJK_CLOSEST_POINTS) return false;
    V3<T> point1 = r.p1, point2_1 = r.p2, normal1 = r.n;
    out.dist = dot(point2_1 - point1, normal1);
    out.point1 = point1;
    out.point2 = inv_rotate(pos12.q, point2_1 - pos12.t);
    out.normal1 = normal1;
    out.normal2 = inv_rotate(pos12.q, -normal1);
    return true;
}

// contact_ball_ball (contact_ball_ball.rs:9-42).
template <class T> bool contact_ball_ball(const Iso<T> &pos12, T r1, T r2, T prediction, Contact<T> &out) {
    V3<T> c = pos12.t;
    T d2 = dot(c, c);
    T sum = r1 + r2;
    T sum_err = sum + prediction;
    if (!(d2 < sum_err * sum_err)) return false;
    V3<T> n1 = d2 != T(0) ? c / norm(c) : v3(T(1), T(0), T(0));
    V3<T> n2 = -inv_rotate(pos12.q, n1);
    out.point1 = n1 * r1;
    out.point2 = n2 * r2;
    out.normal1 = n1;
    out.normal2 = n2;
    out.dist = std::sqrt(d2) - sum;
    return true;
}

// ---- Ball against a convex shape (contact_ball_convex_polyhedron.rs:12-63) ----
// Point projections with solid = false and the feature ids the contact needs.

// Isometry::inverse: (q^-1, q^-1 * -t).
template <class T> inline Iso<T> iso_inverse(const Iso<T> &m) {
    Quat<T> qi = conj(m.q);
    return Iso<T>{qi, rotate(qi, -m.t)};
}

// approx relative_eq!(a, b) with epsilon = max_relative = EPS, per component.
template <class T> inline bool rel_eq(T a, T b) {
    if (a == b) return true;
    if (std::isinf(a) || std::isinf(b)) return false;
    const T e = Eps<T>::eps;
    T d = std::fabs(a - b);
    if (d <= e) return true;
    T fa = std::fabs(a), fb = std::fabs(b);
    T largest = fb > fa ? fb : fa;
    return d <= largest * e;
}
template <class T> inline bool rel_eq(V3<T> a, V3<T> b) { return rel_eq(a.x, b.x) && rel_eq(a.y, b.y) && rel_eq(a.z, b.z); }

// nalgebra Unit::try_new_and_get(v, m): |v|^2 > m^2.
template <class T> inline bool try_unit(V3<T> v, T m, V3<T> &u, T &len) {
    T sq = dot(v, v);
    if (!(sq > m * m)) return false;
    len = std::sqrt(sq);
    u = v / len;
    return true;
}

template <class T> struct PointProj {
    bool inside;
    V3<T> point;
};
// Cuboid feature of Aabb::project_local_point_and_get_feature (point_aabb.rs:60-132).
enum CuboidFeature : uint32_t { CF_UNKNOWN = 0, CF_VERTEX = 1, CF_EDGE = 2, CF_FACE = 3 };

// Segment::project_local_point (point_segment.rs:45-85) with its relative_eq inside flag.
template <class T> PointProj<T> segment_point(V3<T> a, V3<T> b, V3<T> pt) {
    Loc<T> loc;
    V3<T> p = project_segment(a, b, pt, loc);
    return PointProj<T>{rel_eq(p, pt), p};
}

// Cuboid (point_cuboid.rs:14-22) -> Aabb(-he, he) (point_aabb.rs:9-132), solid = false.
template <class T> PointProj<T> cuboid_point_feature(const Shape<T> &s, V3<T> pt, uint32_t &fkind, uint32_t &fid) {
    V3<T> mins = v3(-s.p[0], -s.p[1], -s.p[2]), maxs = v3(s.p[0], s.p[1], s.p[2]);
    V3<T> mins_pt = mins - pt, pt_maxs = pt - maxs;
    V3<T> zero = v3(T(0), T(0), T(0));
    V3<T> shift = vmax(mins_pt, zero) - vmax(pt_maxs, zero);
    bool inside = shift.x == T(0) && shift.y == T(0) && shift.z == T(0);
    V3<T> ls;
    if (!inside) {
        ls = pt + shift;
    } else {
        T best = -std::numeric_limits<T>::max();
        bool is_mins = false;
        int best_id = 0;
        for (int i = 0; i < 3; ++i) {
            T a = mins_pt[i], b = pt_maxs[i];
            if (a < b) {
                if (pt_maxs[i] > best) {
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
        shift[best_id] = is_mins ? best : -best;
        ls = pt + shift;
    }
    int nzero = 0, last_zero = 0, last_nz = 0;
    for (int i = 0; i < 3; ++i) {
        if (shift[i] == T(0)) {
            ++nzero;
            last_zero = i;
        } else {
            last_nz = i;
        }
    }
    const T e = Eps<T>::eps;
    V3<T> center = (mins + maxs) * T(0.5);
    fkind = CF_UNKNOWN;
    fid = 0;
    if (nzero == 3) {
        for (int i = 0; i < 3; ++i) {
            if (ls[i] > maxs[i] - e) {
                fkind = CF_FACE;
                fid = (uint32_t)i;
                break;
            }
            if (ls[i] <= mins[i] + e) {
                fkind = CF_FACE;
                fid = (uint32_t)(i + 3);
                break;
            }
        }
    } else if (nzero == 2) {
        fkind = CF_FACE;
        fid = ls[last_nz] < center[last_nz] ? (uint32_t)(last_nz + 3) : (uint32_t)last_nz;
    } else {
        uint32_t id = 0;
        for (int i = 0; i < 3; ++i)
            if (ls[i] < center[i]) id |= 1u << i;
        if (nzero == 0) {
            fkind = CF_VERTEX;
            fid = id;
        } else {
            fkind = CF_EDGE;
            fid = (id << 2) | (uint32_t)last_zero;
        }
    }
    return PointProj<T>{inside, ls};
}

// Cuboid::feature_normal (cuboid.rs:401-448).
template <class T> bool cuboid_feature_normal(uint32_t fkind, uint32_t id, V3<T> &n) {
    n = v3(T(0), T(0), T(0));
    if (fkind == CF_FACE) {
        if (id < 3) n[id] = T(1);
        else n[id - 3] = T(-1);
        return true;
    }
    if (fkind == CF_EDGE) {
        uint32_t edge = id & 3u, f1 = (edge + 1) % 3, f2 = (edge + 2) % 3, signs = id >> 2;
        n[f1] = (signs & (1u << f1)) ? T(-1) : T(1);
        n[f2] = (signs & (1u << f2)) ? T(-1) : T(1);
        n = n / norm(n);
        return true;
    }
    if (fkind == CF_VERTEX) {
        for (int i = 0; i < 3; ++i) n[i] = (id & (1u << i)) ? T(-1) : T(1);
        n = n / norm(n);
        return true;
    }
    return false;
}

// Capsule::project_local_point (point_capsule.rs:9-50), solid = false.
template <class T> PointProj<T> capsule_point(const Shape<T> &s, V3<T> pt) {
    V3<T> a = v3(T(0), -s.p[0], T(0)), b = v3(T(0), s.p[0], T(0));
    const T r = s.p[1];
    PointProj<T> proj = segment_point(a, b, pt);
    V3<T> dproj = pt - proj.point;
    V3<T> dir;
    T dist;
    if (try_unit(dproj, Eps<T>::eps, dir, dist)) return PointProj<T>{dist <= r, proj.point + dir * r};
    V3<T> sd;
    T sl;
    if (try_unit(b - a, Eps<T>::eps, sd, sl)) {
        // WBasis::orthonormal_basis()[0] (utils/wops.rs:96-110)
        T sign = std::copysign(T(1), sd.z);
        T aa = T(-1) / (sign + sd.z);
        T bb = sd.x * sd.y * aa;
        V3<T> e0 = v3(T(1) + sign * sd.x * sd.x * aa, sign * bb, -sign * sd.x);
        return PointProj<T>{true, proj.point + e0 * r};
    }
    return PointProj<T>{true, proj.point + v3(T(0), r, T(0))};
}

// Cone::project_local_point (point_cone.rs:8-63), solid = false.
template <class T> PointProj<T> cone_point(const Shape<T> &s, V3<T> pt) {
    const T h = s.p[0], r = s.p[1];
    T dx = pt.x, dz = pt.z;
    T planar = std::sqrt(dx * dx + dz * dz);  // Vector2::normalize_mut
    dx = dx / planar;
    dz = dz / planar;
    if (planar <= Eps<T>::eps) {
        dx = T(1);
        dz = T(0);
    }
    V3<T> on_basis = v3(pt.x, -h, pt.z);
    if (pt.y < -h && planar <= r) return PointProj<T>{false, on_basis};
    V3<T> circle = v3(dx * r, -h, dz * r);
    V3<T> apex = v3(T(0), h, T(0));
    V3<T> side_dir = circle - apex;
    PointProj<T> proj = segment_point(apex, circle, pt);
    V3<T> apex_to_basis = v3(T(0), T(-2) * h, T(0));
    if (pt.y >= -h && pt.y <= h && dot(cross(side_dir, pt - apex), cross(side_dir, apex_to_basis)) >= T(0)) {
        V3<T> d1 = proj.point - pt, d2 = on_basis - pt;
        if (dot(d1, d1) > dot(d2, d2)) return PointProj<T>{true, on_basis};
        proj.inside = true;
        return proj;
    }
    return proj;
}

// local_point_projection_on_support_map (point_support_map.rs:17-51), solid = false.
// Returns false where parry reaches unreachable!() (GJK ending without a result).
template <class T> bool support_map_point(const Shape<T> &s, V3<T> pt, PointProj<T> &out, ContactStats *st) {
    Shape<T> origin{ORIGIN, nullptr, 0, {T(0), T(0), T(0), T(0)}};
    const Quat<T> id{T(0), T(0), T(0), T(1)};
    Iso<T> m{id, -pt}, m_inv{id, pt};
    V3<T> dir;
    T len;
    if (!try_unit(pt, Eps<T>::eps, dir, len)) dir = v3(T(1), T(0), T(0));
    Simplex<T> simplex;
    simplex.reset(cso_from_shapes(m_inv, s, origin, dir));
    Iso<T> pos12 = iso_inverse(m);
    uint32_t gi = 0;
    GjkResult<T> g = gjk_closest_points(pos12, s, origin, std::numeric_limits<T>::max(), simplex, &gi);
    if (st) st->gjk_iters = gi;
    if (g.kind == GJK_CLOSEST_POINTS) {
        out = PointProj<T>{false, g.p1};
        return true;
    }
    if (g.kind != GJK_INTERSECTION) return false;
    Epa<T> epa;
    V3<T> p1, p2, n;
    uint32_t ei = 0;
    if (st) st->used_epa = true;
    bool ok = epa.closest_points(pos12, s, origin, simplex, p1, p2, n, &ei);
    if (st) {
        st->epa_iters = ei;
        st->epa_exit = epa.exit;
    }
    out = PointProj<T>{true, ok ? p1 : pt};
    return true;
}

// contact_convex_polyhedron_ball (contact_ball_convex_polyhedron.rs:27-63) in
// shape 1's frame.  Returns 1 (contact), 0 (None), -1 (unreachable!() in parry).
// Triangle::project_local_point_and_get_location (point_triangle.rs:54-291),
// 3D, for any point.  *inside is PointProjection::is_inside: relative_eq!
// (projection, point) in the Voronoi-region branches, true in the solid and
// closest-edge fallbacks of a degenerate face.
template <class T>
V3<T> triangle_project_point(V3<T> a, V3<T> b, V3<T> c, V3<T> pt, bool solid, bool *inside) {
    V3<T> ab = b - a, ac = c - a, ap = pt - a;
    T ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
    auto res = [&](V3<T> proj) {
        *inside = rel_eq(proj, pt);
        return proj;
    };
    if (ab_ap <= T(0) && ac_ap <= T(0)) return res(a);
    V3<T> bp = pt - b;
    T ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
    if (ab_bp >= T(0) && ac_bp <= ab_bp) return res(b);
    V3<T> cp = pt - c;
    T ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
    if (ac_cp >= T(0) && ab_cp <= ac_cp) return res(c);
    V3<T> bc = c - b;
    // stable_check_edges_voronoi (dim3)
    V3<T> n = cross(ab, ac);
    T vc = dot(n, cross(ab, ap));
    if (vc < T(0) && ab_ap >= T(0) && ab_bp <= T(0)) return res(a + ab * (ab_ap / dot(ab, ab)));
    T vb = -dot(n, cross(ac, cp));
    if (vb < T(0) && ac_ap >= T(0) && ac_cp <= T(0)) return res(a + ac * (ac_ap / dot(ac, ac)));
    T va = dot(n, cross(bc, bp));
    if (va < T(0) && ac_bp - ab_bp >= T(0) && ab_cp - ac_cp >= T(0)) return res(b + bc * (dot(bc, bp) / dot(bc, bc)));
    if (va + vb + vc != T(0)) {
        T denom = T(1) / (va + vb + vc);
        T v = vb * denom;
        T w = vc * denom;
        return res((a + ab * v) + ac * w);
    }
    *inside = true;
    if (solid) return pt;
    // Degenerate face, not solid: the closest of the three edges.
    T v = ab_ap / (ab_ap - ab_bp);
    T w = ac_ap / (ac_ap - ac_cp);
    T u = (ac_bp - ab_bp) / (ac_bp - ab_bp + ab_cp - ac_cp);
    T d_ab = dot(ap, ap) - (dot(ab, ab) * v * v);
    T d_ac = dot(ap, ap) - (dot(ac, ac) * w * w);
    T d_bc = dot(bp, bp) - (dot(bc, bc) * u * u);
    if (d_ab < d_ac) return d_ab < d_bc ? a + ab * v : b + bc * u;
    return d_ac < d_bc ? a + ac * w : b + bc * u;
}

template <class T>
int contact_convex_ball(const Iso<T> &pos12, const Shape<T> &g1, T radius, T prediction, Contact<T> &out,
                        ContactStats *st) {
    V3<T> c = pos12.t;
    PointProj<T> proj;
    uint32_t fkind = CF_UNKNOWN, fid = 0;
    switch (g1.type) {
        case CUBOID: proj = cuboid_point_feature(g1, c, fkind, fid); break;
        case CAPSULE: proj = capsule_point(g1, c); break;
        case CONE: proj = cone_point(g1, c); break;
        case CONVEX_POLYHEDRON:
            if (!support_map_point(g1, c, proj, st)) return -1;
            break;
        case TRIANGLE: {  // project_local_point_and_get_feature: solid = true in 3D (point_triangle.rs:26-48)
            bool inside = false;
            proj.point = triangle_project_point(g1.pts[0], g1.pts[1], g1.pts[2], c, true, &inside);
            proj.inside = inside;
            break;
        }
        default: return -1;
    }
    T dist;
    V3<T> n1, dir1;
    T len;
    if (try_unit(proj.point - c, Eps<T>::eps, dir1, len)) {
        if (proj.inside) {
            dist = -len - radius;
            n1 = dir1;
        } else {
            dist = len - radius;
            n1 = -dir1;
        }
    } else {
        // A Cuboid's feature normal; a Triangle's normal() for any feature
        // (triangle.rs:226-228, 626-628; shape.rs:919-928).  A ConvexPolyhedron's
        // feature is Unknown whenever this branch is taken (point_support_map.rs:66-72
        // uses the same threshold), Capsule and Cone have none (shape.rs:425-431).
        dist = -radius;
        V3<T> u;
        T l;
        bool have = g1.type == CUBOID && cuboid_feature_normal(fkind, fid, n1);
        if (g1.type == TRIANGLE && try_unit(cross(g1.pts[1] - g1.pts[0], g1.pts[2] - g1.pts[0]), Eps<T>::eps, u, l)) {
            n1 = u;
            have = true;
        }
        if (!have) {
            if (try_unit(proj.point, Eps<T>::eps, u, l)) n1 = u;
            else n1 = v3(T(0), T(1), T(0));
        }
    }
    if (!(dist <= prediction)) return 0;
    V3<T> n2 = inv_rotate(pos12.q, -n1);
    out.point1 = proj.point;
    out.point2 = n2 * radius;
    out.normal1 = n1;
    out.normal2 = n2;
    out.dist = dist;
    return 1;
}

// query::contact (contact_shape_shape.rs:123-138) with the dispatch of
// DefaultQueryDispatcher::contact (default_query_dispatcher.rs:302-356) for the
// shapes held here.  Returns 1 on contact, 0 for None, -1 where parry reaches
// unreachable!().
template <class T>
int contact(const Iso<T> &pos1, const Shape<T> &g1, const Iso<T> &pos2, const Shape<T> &g2, T prediction,
            Contact<T> &out, const V3<T> *init_dir, ContactStats *st) {
    Iso<T> pos12 = inv_mul(pos1, pos2);
    bool ok;
    if (g1.type == BALL && g2.type == BALL) {
        ok = contact_ball_ball(pos12, g1.p[0], g2.p[0], prediction, out);
    } else if (g1.type == BALL) {
        // contact_ball_convex_polyhedron: the flipped contact of (shape 2, ball) under pos12^-1
        int r = contact_convex_ball(iso_inverse(pos12), g2, g1.p[0], prediction, out, st);
        if (r < 0) return -1;
        ok = r == 1;
        if (ok) {
            std::swap(out.point1, out.point2);
            std::swap(out.normal1, out.normal2);
        }
    } else if (g2.type == BALL) {
        int r = contact_convex_ball(pos12, g1, g2.p[0], prediction, out, st);
        if (r < 0) return -1;
        ok = r == 1;
    } else {
        ok = contact_sm_sm(pos12, g1, g2, prediction, out, init_dir, st);
    }
    if (!ok) return 0;
    // Contact::transform_by_mut (contact.rs:171-176)
    out.point1 = transform_point(pos1, out.point1);
    out.point2 = transform_point(pos2, out.point2);
    out.normal1 = rotate(pos1.q, out.normal1);
    out.normal2 = rotate(pos2.q, out.normal2);
    return 1;
}

}  // namespace oracle
