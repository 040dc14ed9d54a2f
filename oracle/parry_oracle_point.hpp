// parry_oracle_point.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of more TriMesh queries of parry3d, for the GPU parity
// tests: point projection through the Bvh (TriMesh::project_local_point,
// src/query/point/point_composite_shape.rs:59-80 and 164-168), the ray cast
// with back- or front-face culling (TriMesh::cast_local_ray_with_culling,
// src/query/ray/ray_trimesh.rs:58-65, 131-178; in parry_oracle.hpp), and the
// contact between a TriMesh and a convex shape (contact_composite_shape_shape,
// src/query/contact/contact_composite_shape_shape.rs:14-59).  Same conventions as
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

// Bvh::leaves with an AABB-intersection check (bvh_traverse.rs:1-68, through
// Bvh::intersect_aabb, bvh_queries.rs:203-205): leaves in depth-first order,
// left subtree before right, a subtree skipped when its box misses.
template <class T, class F> void bvh_leaves_intersecting(const Bvh<T> &bvh, const Aabb<T> &box, F &&emit) {
    const auto &nodes = bvh.nodes;
    if (nodes.empty()) return;
    std::vector<const Node<T> *> stack;
    const Node<T> *next = nullptr;
    if (nodes[0].left.aabb().intersects(box)) next = &nodes[0].left;
    if (nodes[0].right.leaf_count() > 0 && nodes[0].right.aabb().intersects(box)) stack.push_back(&nodes[0].right);
    for (;;) {
        if (!next) {
            if (stack.empty()) return;
            next = stack.back();
            stack.pop_back();
        }
        const Node<T> *node = next;
        next = nullptr;
        if (node->is_leaf()) {
            emit(node->children);
            continue;
        }
        const Wide<T> &ch = nodes[node->children];
        if (ch.left.aabb().intersects(box)) next = &ch.left;
        if (ch.right.aabb().intersects(box)) {
            if (!next) next = &ch.right;
            else stack.push_back(&ch.right);
        }
    }
}

// query::contact(pos1, TriMesh, pos2, g2) for a convex g2: the composite
// dispatch (default_query_dispatcher.rs:343-346) to
// CompositeShapeRef::contact_with_shape (contact_composite_shape_shape.rs:14-42).
// g2's AABB at pos12 loosened by `prediction` selects the triangles; each goes
// through the contact dispatch as a Triangle with pos12 (a TriMesh part has no
// pose of its own); the smallest dist wins, the first on ties; an Err or None
// is skipped.  The result is in world space, as query::contact returns it.
// Returns 1 with *tri set, or 0.
template <class T>
int trimesh_contact_shape(const TriMesh<T> &m, const Iso<T> &pos1, const Shape<T> &g2, const Iso<T> &pos2,
                          T prediction, Contact<T> &out, uint32_t *tri, uint64_t *candidates) {
    Iso<T> pos12 = inv_mul(pos1, pos2);
    Aabb<T> box = shape_aabb(pos12, g2);
    box.mins = box.mins + v3(-prediction, -prediction, -prediction);  // Aabb::loosened
    box.maxs = box.maxs + v3(prediction, prediction, prediction);
    int have = 0;
    bvh_leaves_intersecting(m.bvh, box, [&](uint32_t prim) {
        if (candidates) ++*candidates;
        V3<T> p[3];
        m.tri(prim, p[0], p[1], p[2]);
        Shape<T> t{TRIANGLE, p, 3, {T(0), T(0), T(0), T(0)}};
        Contact<T> c{};
        ContactStats st;
        if (contact(pos1, t, pos2, g2, prediction, c, static_cast<const V3<T> *>(nullptr), &st) != 1) return;
        if (!have || c.dist < out.dist) {
            out = c;
            *tri = prim;
            have = 1;
        }
    });
    return have;
}

}  // namespace oracle
