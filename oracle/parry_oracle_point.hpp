// parry_oracle_point.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of two more TriMesh queries of parry3d, for the GPU parity
// tests: point projection through the Bvh (TriMesh::project_local_point,
// src/query/point/point_composite_shape.rs:59-80 and 164-168) and the ray
// cast with back- or front-face culling (TriMesh::cast_local_ray_with_culling,
// src/query/ray/ray_trimesh.rs:58-65, 131-178).  Same conventions as
// parry_oracle.hpp: T = float is parry3d, T = double parry3d-f64, and every
// expression keeps the reference's operation order.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "parry_oracle.hpp"
#include "parry_oracle_contact.hpp"

namespace oracle {

// Aabb::distance_to_local_point with solid = true (point_aabb.rs:135-146):
// |max(mins - p, p - maxs, 0)|.
template <class T> inline T aabb_point_distance(const Aabb<T> &b, V3<T> p) {
    return norm(vmax(vmax(b.mins - p, p - b.maxs), v3(T(0), T(0), T(0))));
}

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

template <class T> struct PointHit {
    V3<T> point;
    bool inside;
    uint32_t tri;
};

// TriMesh::project_local_point (point_composite_shape.rs:59-80, 164-168):
// Bvh::find_best (bvh_traverse.rs:335-417) from max cost Real::MAX, node cost
// the AABB distance, leaf cost |p - projection| (na::distance).
template <class T>
bool trimesh_project_point(const TriMesh<T> &m, V3<T> pt, bool solid, PointHit<T> *out,
                           RayCounters *cnt = nullptr) {
    const auto &nodes = m.bvh.nodes;
    if (nodes.empty()) return false;
    auto leaf_cost = [&](uint32_t prim, PointHit<T> *h) -> T {
        if (cnt) cnt->leaf_tests++;
        V3<T> a, b, c;
        m.tri(prim, a, b, c);
        bool ins = false;
        V3<T> p = triangle_project_point(a, b, c, pt, solid, &ins);
        *h = PointHit<T>{p, ins, prim};
        return norm(pt - p);
    };
    const T max_cost = Consts<T>::max_value;
    if (nodes[0].right.leaf_count() == 0) {  // partial root, bvh_traverse.rs:348-359
        if (cnt) cnt->wide_loads++;
        const Node<T> &leaf = nodes[0].left;
        if (!(aabb_point_distance(leaf.aabb(), pt) < max_cost)) return false;
        PointHit<T> h;
        if (!(leaf_cost(leaf.children, &h) < max_cost)) return false;
        *out = h;
        return true;
    }
    std::vector<uint32_t> stack;
    T best_cost = max_cost;
    bool have = false;
    PointHit<T> best{};
    uint32_t curr = 0;
    for (;;) {
        if (cnt) cnt->wide_loads++;
        const Wide<T> &w = nodes[curr];
        const Node<T> *left = &w.left;
        const Node<T> *right = &w.right;
        T ls = aabb_point_distance(left->aabb(), pt);
        T rs = aabb_point_distance(right->aabb(), pt);
        if (ls > rs) {
            std::swap(ls, rs);
            std::swap(left, right);
        }
        bool found_next = false;
        if (ls < best_cost && ls != max_cost) {
            if (left->is_leaf()) {
                PointHit<T> h;
                T cst = leaf_cost(left->children, &h);
                if (cst < best_cost) {
                    best = h;
                    best_cost = cst;
                    have = true;
                }
            } else {
                curr = left->children;
                found_next = true;
            }
        }
        if (rs < best_cost && rs != max_cost) {
            if (right->is_leaf()) {
                PointHit<T> h;
                T cst = leaf_cost(right->children, &h);
                if (cst < best_cost) {
                    best = h;
                    best_cost = cst;
                    have = true;
                }
            } else if (found_next) {
                stack.push_back(right->children);
            } else {
                curr = right->children;
                found_next = true;
            }
        }
        if (!found_next) {
            if (stack.empty()) break;
            curr = stack.back();
            stack.pop_back();
        }
    }
    if (have) *out = best;
    return have;
}

// Brute force over all triangles: the projection at the smallest distance
// (first index on ties).  Test helper: pins the Bvh traversal above.
template <class T> bool trimesh_project_point_brute(const TriMesh<T> &m, V3<T> pt, bool solid, PointHit<T> *out) {
    bool have = false;
    T best = Consts<T>::max_value;
    for (uint32_t i = 0; i < m.num_triangles(); ++i) {
        V3<T> a, b, c;
        m.tri(i, a, b, c);
        bool ins = false;
        V3<T> p = triangle_project_point(a, b, c, pt, solid, &ins);
        T d = norm(pt - p);
        if (d < best) {
            best = d;
            *out = PointHit<T>{p, ins, i};
            have = true;
        }
    }
    return have;
}

}  // namespace oracle
