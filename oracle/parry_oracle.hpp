// parry_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of parry3d's batched collision-query hot path, used as the
// correctness oracle for the gfx950 kernels and as the CPU baseline timed by
// bench.py.  Nothing in the product library links or calls this file: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
// shared object built from it (oracle/_build/liboracle.so).
//
// Every function is a template over the scalar T so that T=float restates
// parry3d (f32) and T=double restates parry3d-f64 (same Rust source compiled
// with Real=f64, src/lib.rs:81-89).  The f32 instantiation must reproduce the
// reference's floating-point operation sequence exactly, so this file is
// compiled with -ffp-contract=off and without fast-math; every citation below
// is a path:line in dimforge/parry 0.25.3.
//
// nalgebra 0.34 arithmetic that parry relies on (not present in the reference
// tree) is restated from its published implementation:
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx)
//   norm(v)       = sqrt(dot(v,v)); normalize(v) = v / norm(v) per component
//   center(a,b)   = (a + b) * 0.5
//   inf/sup       = per-component f32::min / f32::max (NaN-ignoring, = fmin/fmax)
//   imax(v)       = index of the first strict maximum
//
// Parity status (how each function is pinned) is listed in the header comment
// of each section and checked by tests/test_oracle_*.py.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

template <class T> struct Consts;
template <> struct Consts<float> {
    static constexpr float bin_eps = 1.0e-5f;        // bvh_binned_build.rs:43
    static constexpr float max_value = 3.40282347e+38f;
};
template <> struct Consts<double> {
    static constexpr double bin_eps = 1.0e-5;
    static constexpr double max_value = 1.7976931348623157e+308;
};

template <class T> struct V3 {
    T x, y, z;
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a / norm(a); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) {
    return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) {
    return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}
template <class T> inline int imax(V3<T> v) {
    int ip = 0;
    T m = v.x;
    if (v.y > m) { m = v.y; ip = 1; }
    if (v.z > m) { ip = 2; }
    return ip;
}

// Rust `f as usize` (saturating, NaN -> 0).  Only the values 0..=7 are reachable
// without a Rust panic in bvh_binned_build.rs:57, larger ones are clamped to
// u64::MAX as the language does.
template <class T> inline uint64_t rust_as_usize(T f) {
    if (!(f > T(0))) return 0;  // NaN, negative, zero
    if (f >= T(18446744073709551616.0)) return UINT64_MAX;
    return (uint64_t)f;
}

// ---------------------------------------------------------------------------
// Aabb (src/bounding_volume/aabb.rs)
// ---------------------------------------------------------------------------
template <class T> struct Aabb {
    V3<T> mins, maxs;
    static Aabb invalid() {  // aabb.rs:268-273
        T m = Consts<T>::max_value;
        return {v3(m, m, m), v3(-m, -m, -m)};
    }
    V3<T> center() const { return (mins + maxs) * T(0.5); }  // aabb.rs:392 via na::center
    V3<T> extents() const { return maxs - mins; }              // aabb.rs:480
    T volume() const {                                          // aabb.rs:448-454
        V3<T> e = extents();
        return e.x * e.y * e.z;
    }
    void merge(const Aabb &o) {  // aabb.rs:961-964
        mins = vmin(mins, o.mins);
        maxs = vmax(maxs, o.maxs);
    }
    bool intersects(const Aabb &o) const {  // aabb.rs:951-953 (partial_le / partial_ge)
        return mins.x <= o.maxs.x && mins.y <= o.maxs.y && mins.z <= o.maxs.z &&
               maxs.x >= o.mins.x && maxs.y >= o.mins.y && maxs.z >= o.mins.z;
    }
};

// Triangle::local_aabb (aabb_triangle.rs:16-30): min(a,b) then min(.,c).
template <class T> inline Aabb<T> triangle_aabb(V3<T> a, V3<T> b, V3<T> c) {
    Aabb<T> r;
    for (int d = 0; d < 3; ++d) {
        r.mins[d] = std::fmin(std::fmin(a[d], b[d]), c[d]);
        r.maxs[d] = std::fmax(std::fmax(a[d], b[d]), c[d]);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Bvh data model (src/partitioning/bvh/bvh_tree.rs)
// ---------------------------------------------------------------------------
static constexpr uint32_t CHANGED = 0b01;         // bvh_tree.rs:157
static constexpr uint32_t CHANGE_PENDING = 0b11;  // bvh_tree.rs:158

template <class T> struct Node {  // BvhNode, bvh_tree.rs:456-467
    V3<T> mins;
    uint32_t children;
    V3<T> maxs;
    uint32_t data;  // leaf_count (30 bits) | change flags << 30
    uint32_t leaf_count() const { return data & 0x3fffffffu; }
    bool is_leaf() const { return leaf_count() == 1; }
    Aabb<T> aabb() const { return {mins, maxs}; }
    V3<T> center() const { return (mins + maxs) * T(0.5); }
    static Node zeros() { return Node{v3(T(0), T(0), T(0)), 0, v3(T(0), T(0), T(0)), 0}; }
    static Node leaf(const Aabb<T> &a, uint32_t id) {  // bvh_tree.rs:520-527
        return Node{a.mins, id, a.maxs, 1u | (CHANGE_PENDING << 30)};
    }
    void resolve_pending_change() {  // bvh_tree.rs:192-198
        if ((data >> 30) == CHANGE_PENDING)
            data = (data & 0x3fffffffu) | (CHANGED << 30);
        else
            data = data & 0x3fffffffu;
    }
    Node merged(const Node &o, uint32_t children_id) const {  // bvh_tree.rs:610-618, 200-204
        uint32_t lc = leaf_count() + o.leaf_count();
        uint32_t ch = (data >> 30) | (o.data >> 30);
        return Node{vmin(mins, o.mins), children_id, vmax(maxs, o.maxs), lc | (ch << 30)};
    }
    bool intersects(const Node &o) const {  // bvh_tree.rs:954-957
        return aabb().intersects(o.aabb());
    }
};

template <class T> struct Wide {  // BvhNodeWide, bvh_tree.rs:263-266
    Node<T> left, right;
    Node<T> merged(uint32_t my_id) const { return left.merged(right, my_id); }
    uint32_t leaf_count() const { return left.leaf_count() + right.leaf_count(); }
};

template <class T> struct Bvh {
    std::vector<Wide<T>> nodes;
    uint32_t leaf_count() const { return nodes.empty() ? 0 : nodes[0].leaf_count(); }
};

// rebuild_range_binned (bvh_binned_build.rs:39-176).  Topology is identical to
// the reference: same bins, same volume*count cost, same in-place swap
// partition (the leaf order inside a range matters when a split degenerates
// to len/2).  Nodes are appended in the same order as the reference, which is
// depth-first preorder.
template <class T> struct BinnedBuilder {
    std::vector<Wide<T>> &nodes;

    struct Bin {
        Aabb<T> aabb;
        uint32_t leaf_count;
    };

    static uint64_t bin_of(const Node<T> &n, int axis, T k0, T k1) {
        return rust_as_usize(k1 * (n.center()[axis] - k0));
    }

    void build(uint32_t target, Node<T> *leaves, size_t len) {
        const int NUM_BINS = 8;
        const T BIN_EPSILON = Consts<T>::bin_eps;
        Bin bins[NUM_BINS];
        for (auto &b : bins) b = Bin{Aabb<T>::invalid(), 0};

        // Aabb::from_points over leaf centres (aabb_utils.rs local_point_cloud_aabb).
        V3<T> c0 = leaves[0].center();
        Aabb<T> cent{c0, c0};
        for (size_t i = 1; i < len; ++i) {
            V3<T> c = leaves[i].center();
            cent.mins = vmin(cent.mins, c);
            cent.maxs = vmax(cent.maxs, c);
        }
        int axis = imax(cent.extents());
        T r0 = cent.mins[axis], r1 = cent.maxs[axis];
        T k1 = T(NUM_BINS) * (T(1) - BIN_EPSILON) / (r1 - r0);
        T k0 = r0;
        for (size_t i = 0; i < len; ++i) {
            uint64_t b = bin_of(leaves[i], axis, k0, k1);
            if (b >= (uint64_t)NUM_BINS) b = NUM_BINS - 1;  // unreachable in Rust without panic
            bins[b].aabb.merge(leaves[i].aabb());
            bins[b].leaf_count += 1;
        }

        Bin right_merges[NUM_BINS];
        for (int i = 0; i < NUM_BINS; ++i) right_merges[i] = bins[i];
        Bin right_acc = bins[NUM_BINS - 1];
        for (int i = 1; i < NUM_BINS - 1; ++i) {
            right_acc.aabb.merge(right_merges[NUM_BINS - 1 - i].aabb);
            right_acc.leaf_count += right_merges[NUM_BINS - 1 - i].leaf_count;
            right_merges[NUM_BINS - 1 - i] = right_acc;
        }

        T best_cost = Consts<T>::max_value;
        int best_plane = 0;
        Bin left_merge = bins[0];
        uint32_t best_leaf_count = bins[0].leaf_count;
        for (int i = 0; i < NUM_BINS - 1; ++i) {
            const Bin &right = right_merges[i + 1];
            T cost = left_merge.aabb.volume() * (T)left_merge.leaf_count +
                     right.aabb.volume() * (T)right.leaf_count;
            if (cost < best_cost) {
                best_cost = cost;
                best_plane = i;
                best_leaf_count = left_merge.leaf_count;
            }
            left_merge.aabb.merge(bins[i + 1].aabb);
            left_merge.leaf_count += bins[i + 1].leaf_count;
        }

        size_t mid = best_leaf_count;
        if (mid == 0 || mid == len) {
            mid = len / 2;
        } else {
            size_t l = 0, r = mid;
            while (l != mid && r != len) {
                bool stop = false;
                while (bin_of(leaves[l], axis, k0, k1) <= (uint64_t)best_plane) {
                    ++l;
                    if (l == mid) { stop = true; break; }
                }
                if (stop) break;
                while (bin_of(leaves[r], axis, k0, k1) > (uint64_t)best_plane) {
                    ++r;
                    if (r == len) { stop = true; break; }
                }
                if (stop) break;
                std::swap(leaves[l], leaves[r]);
                ++l;
                ++r;
            }
        }

        // Left child.
        if (mid == 1) {
            nodes[target].left = leaves[0];
        } else {
            uint32_t lid = (uint32_t)nodes.size();
            nodes.push_back(Wide<T>{Node<T>::zeros(), Node<T>::zeros()});
            build(lid, leaves, mid);
            nodes[target].left = nodes[lid].merged(lid);
        }
        // Right child.
        if (len - mid == 1) {
            nodes[target].right = leaves[mid];
        } else {
            uint32_t rid = (uint32_t)nodes.size();
            nodes.push_back(Wide<T>{Node<T>::zeros(), Node<T>::zeros()});
            build(rid, leaves + mid, len - mid);
            nodes[target].right = nodes[rid].merged(rid);
        }
    }
};

// refit with depth-first relayout (bvh_refit.rs:182-320).
template <class T> struct Refitter {
    const std::vector<Wide<T>> &src;
    std::vector<Wide<T>> &dst;
    uint32_t len;

    // parent: (wide index, is_right)
    void recurse(uint32_t source_id, uint32_t parent_id, bool parent_right) {
        uint32_t target_id = len++;
        const Wide<T> &n = src[source_id];
        if (!n.left.is_leaf()) {
            recurse(n.left.children, target_id, false);
        } else {
            dst[target_id].left = n.left;
            dst[target_id].left.resolve_pending_change();
        }
        if (!n.right.is_leaf()) {
            recurse(n.right.children, target_id, true);
        } else {
            dst[target_id].right = n.right;
            dst[target_id].right.resolve_pending_change();
        }
        const Wide<T> &t = dst[target_id];
        Node<T> m = t.left.merged(t.right, target_id);
        if (parent_right) dst[parent_id].right = m; else dst[parent_id].left = m;
    }
};

template <class T> void refit(Bvh<T> &bvh) {
    std::vector<Wide<T>> &src = bvh.nodes;
    if (src.empty()) return;
    if (src[0].leaf_count() <= 2) {
        Wide<T> root = src[0];
        root.left.resolve_pending_change();
        if (root.right.leaf_count() > 0) root.right.resolve_pending_change();
        src.assign(1, root);
        return;
    }
    std::vector<Wide<T>> dst(src.size(), Wide<T>{Node<T>::zeros(), Node<T>::zeros()});
    Refitter<T> r{src, dst, 1};
    if (!src[0].left.is_leaf()) {
        r.recurse(src[0].left.children, 0, false);
    } else {
        dst[0].left = src[0].left;
        dst[0].left.resolve_pending_change();
    }
    if (!src[0].right.is_leaf()) {
        r.recurse(src[0].right.children, 0, true);
    } else {
        dst[0].right = src[0].right;
        dst[0].right.resolve_pending_change();
    }
    dst.resize(r.len);
    bvh.nodes.swap(dst);
}

// Bvh::from_iter with BvhBuildStrategy::Binned (bvh_tree.rs:1891-1955).
template <class T> Bvh<T> bvh_from_leaves(const std::vector<Aabb<T>> &aabbs) {
    Bvh<T> bvh;
    std::vector<Node<T>> leaves;
    leaves.reserve(aabbs.size());
    for (size_t i = 0; i < aabbs.size(); ++i) leaves.push_back(Node<T>::leaf(aabbs[i], (uint32_t)i));
    if (leaves.empty()) return bvh;
    if (leaves.size() == 1) {
        bvh.nodes.push_back(Wide<T>{leaves[0], Node<T>::zeros()});
        return bvh;
    }
    if (leaves.size() == 2) {
        bvh.nodes.push_back(Wide<T>{leaves[0], leaves[1]});
        return bvh;
    }
    bvh.nodes.reserve(leaves.size());
    bvh.nodes.push_back(Wide<T>{Node<T>::zeros(), Node<T>::zeros()});
    BinnedBuilder<T> b{bvh.nodes};
    b.build(0, leaves.data(), leaves.size());
    refit(bvh);
    return bvh;
}

// ---- PLOC build (bvh_ploc_build.rs:10-93, utils/morton.rs:13-40) --------------

// split_by_3_u64 and morton_encode_u64 (morton.rs:13-31): 21 bits per axis.
inline uint64_t morton_split_by_3(uint32_t a) {
    uint64_t x = (uint64_t)a & 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// Rust's `f64 as u32`: truncation, saturating, NaN -> 0.
inline uint32_t rust_as_u32(double f) {
    if (!(f > 0.0)) return 0u;
    if (f >= 4294967296.0) return UINT32_MAX;
    return (uint32_t)f;
}
// morton_encode_u64_unorm (morton.rs:33-40): p * 2^21, each axis `as u32`.
inline uint64_t morton_encode_unorm(double x, double y, double z) {
    const double s = (double)(1u << 21);
    return morton_split_by_3(rust_as_u32(x * s)) | morton_split_by_3(rust_as_u32(y * s)) << 1 |
           morton_split_by_3(rust_as_u32(z * s)) << 2;
}

// Aabb::half_area (aabb.rs:473-476).
template <class T> inline T half_area(const Aabb<T> &a) {
    V3<T> e = a.maxs - a.mins;
    return e.x * (e.y + e.z) + e.y * e.z;
}

// The Morton key of each leaf (bvh_ploc_build.rs:12-21): its centre relative to
// the centres' AABB, times the inverse extents (in T), widened to f64.
template <class T> std::vector<uint64_t> ploc_keys(const std::vector<Node<T>> &leaves) {
    V3<T> c0 = leaves[0].center();
    Aabb<T> cent{c0, c0};
    for (size_t i = 1; i < leaves.size(); ++i) {
        V3<T> c = leaves[i].center();
        cent.mins = vmin(cent.mins, c);
        cent.maxs = vmax(cent.maxs, c);
    }
    V3<T> e = cent.extents();
    V3<T> inv = v3(T(1) / e.x, T(1) / e.y, T(1) / e.z);
    std::vector<uint64_t> keys(leaves.size());
    for (size_t i = 0; i < leaves.size(); ++i) {
        V3<T> c = leaves[i].center() - cent.mins;
        c = v3(c.x * inv.x, c.y * inv.y, c.z * inv.z);
        keys[i] = morton_encode_unorm((double)c.x, (double)c.y, (double)c.z);
    }
    return keys;
}

// rebuild_range_ploc (bvh_ploc_build.rs:10-93): leaves stably sorted by Morton
// key (sort_by_cached_key), then levels of merges: every node picks, within 16
// places either side, the partner with the smallest half area of the merged box
// (the first on ties); mutual picks merge (in order of the lower index, which
// also takes the merged node's place), the others carry over.
template <class T> void rebuild_ploc(std::vector<Wide<T>> &nodes, uint32_t target, std::vector<Node<T>> leaves) {
    {
        std::vector<uint64_t> keys = ploc_keys(leaves);
        std::vector<size_t> order(leaves.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
        std::vector<Node<T>> sorted(leaves.size());
        for (size_t i = 0; i < order.size(); ++i) sorted[i] = leaves[order[i]];
        leaves.swap(sorted);
    }
    const size_t R = 16;
    std::vector<size_t> cand(leaves.size(), SIZE_MAX);
    std::vector<Node<T>> next;
    next.reserve(leaves.size());
    while (leaves.size() > 1) {
        const size_t L = leaves.size();
        for (size_t i = 0; i < L; ++i) {
            T best_sah = Consts<T>::max_value;
            size_t best = SIZE_MAX;
            const size_t lo = i >= R ? i - R : 0, hi = std::min(i + R, L - 1);
            for (size_t k = lo; k <= hi; ++k) {
                if (k == i) continue;
                Aabb<T> m{vmin(leaves[i].mins, leaves[k].mins), vmax(leaves[i].maxs, leaves[k].maxs)};
                T sah = half_area(m);
                if (sah < best_sah) {
                    best_sah = sah;
                    best = k;
                }
            }
            cand[i] = best;
        }
        for (size_t i = 0; i < L; ++i) {
            const size_t k = cand[i];
            if (cand[k] == i) {
                if (i > k) continue;
                Wide<T> w{leaves[i], leaves[k]};
                if (L == 2) {
                    nodes[target] = w;
                } else {
                    const uint32_t id = (uint32_t)nodes.size();
                    nodes.push_back(w);
                    next.push_back(w.merged(id));
                }
            } else {
                next.push_back(leaves[i]);
            }
        }
        leaves.swap(next);
        next.clear();
    }
}

// Bvh::from_iter with BvhBuildStrategy::Ploc (bvh_tree.rs:1891-1955).
template <class T> Bvh<T> bvh_from_leaves_ploc(const std::vector<Aabb<T>> &aabbs) {
    Bvh<T> bvh;
    std::vector<Node<T>> leaves;
    leaves.reserve(aabbs.size());
    for (size_t i = 0; i < aabbs.size(); ++i) leaves.push_back(Node<T>::leaf(aabbs[i], (uint32_t)i));
    if (leaves.empty()) return bvh;
    if (leaves.size() <= 2) {
        bvh.nodes.push_back(Wide<T>{leaves[0], leaves.size() == 2 ? leaves[1] : Node<T>::zeros()});
        return bvh;
    }
    bvh.nodes.reserve(leaves.size());
    bvh.nodes.push_back(Wide<T>{Node<T>::zeros(), Node<T>::zeros()});
    rebuild_ploc(bvh.nodes, 0, std::move(leaves));
    refit(bvh);
    return bvh;
}

// Depth of the tree in wide-node levels (root = 1).
template <class T> uint32_t bvh_depth(const Bvh<T> &bvh) {
    if (bvh.nodes.empty()) return 0;
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 1u}};
    uint32_t best = 0;
    while (!st.empty()) {
        auto [id, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        const Wide<T> &w = bvh.nodes[id];
        if (w.right.leaf_count() == 0) continue;
        if (!w.left.is_leaf()) st.push_back({w.left.children, d + 1});
        if (!w.right.is_leaf()) st.push_back({w.right.children, d + 1});
    }
    return best;
}

// ---------------------------------------------------------------------------
// Rays (src/query/ray)
// ---------------------------------------------------------------------------
template <class T> struct Ray {
    V3<T> origin, dir;
};

// Aabb::cast_local_ray (ray_aabb.rs:12-49).  Returns false for None.
template <class T> inline bool aabb_cast_local_ray(const Aabb<T> &a, const Ray<T> &ray, T max_toi,
                                                   bool solid, T *out) {
    T tmin = T(0);
    T tmax = max_toi;
    for (int i = 0; i < 3; ++i) {
        if (ray.dir[i] == T(0)) {
            if (ray.origin[i] < a.mins[i] || ray.origin[i] > a.maxs[i]) return false;
        } else {
            T denom = T(1) / ray.dir[i];
            T inter_near = (a.mins[i] - ray.origin[i]) * denom;
            T inter_far = (a.maxs[i] - ray.origin[i]) * denom;
            if (inter_near > inter_far) std::swap(inter_near, inter_far);
            tmin = std::fmax(tmin, inter_near);
            tmax = std::fmin(tmax, inter_far);
            if (tmin > tmax) return false;
        }
    }
    *out = (tmin == T(0) && !solid) ? tmax : tmin;
    return true;
}

// BvhNode::cast_ray (bvh_tree.rs:1177-1181): solid = true, miss -> Real::MAX.
template <class T> inline T node_cast_ray(const Node<T> &n, const Ray<T> &ray, T max_toi) {
    T t;
    if (aabb_cast_local_ray(n.aabb(), ray, max_toi, true, &t)) return t;
    return Consts<T>::max_value;
}

template <class T> struct RayHit {
    T toi;
    V3<T> normal;
    uint32_t fid;  // FeatureId::Face(fid)
};

// local_ray_intersection_with_triangle (ray_triangle.rs:70-152).
template <class T> inline bool ray_triangle(V3<T> a, V3<T> b, V3<T> c, const Ray<T> &ray, RayHit<T> *out) {
    V3<T> ab = b - a;
    V3<T> ac = c - a;
    V3<T> n = cross(ab, ac);
    T d = dot(n, ray.dir);
    if (d == T(0)) return false;
    V3<T> ap = ray.origin - a;
    T t = dot(ap, n);
    if ((t < T(0) && d < T(0)) || (t > T(0) && d > T(0))) return false;
    uint32_t fid = d < T(0) ? 0 : 1;
    d = std::fabs(d);
    V3<T> e = -cross(ray.dir, ap);
    T v, w, toi;
    V3<T> normal;
    if (t < T(0)) {
        v = -dot(ac, e);
        if (v < T(0) || v > d) return false;
        w = dot(ab, e);
        if (w < T(0) || v + w > d) return false;
        T invd = T(1) / d;
        toi = -t * invd;
        normal = -normalize(n);
    } else {
        v = dot(ac, e);
        if (v < T(0) || v > d) return false;
        w = -dot(ab, e);
        if (w < T(0) || v + w > d) return false;
        T invd = T(1) / d;
        toi = t * invd;
        normal = normalize(n);
    }
    *out = RayHit<T>{toi, normal, fid};
    return true;
}

// Triangle::cast_local_ray_and_get_normal (ray_triangle.rs:49-62): `solid` ignored.
template <class T> inline bool triangle_cast(V3<T> a, V3<T> b, V3<T> c, const Ray<T> &ray, T max_toi,
                                             RayHit<T> *out) {
    RayHit<T> h;
    if (!ray_triangle(a, b, c, ray, &h)) return false;
    if (h.toi <= max_toi) {
        *out = h;
        return true;
    }
    return false;
}

template <class T> struct TriMesh {
    std::vector<V3<T>> vertices;
    std::vector<uint32_t> indices;  // 3 per triangle
    Bvh<T> bvh;
    uint32_t num_triangles() const { return (uint32_t)(indices.size() / 3); }
    void tri(uint32_t i, V3<T> &a, V3<T> &b, V3<T> &c) const {  // trimesh.rs:1896-1903
        a = vertices[indices[3 * i]];
        b = vertices[indices[3 * i + 1]];
        c = vertices[indices[3 * i + 2]];
    }
};

// TriMesh::rebuild_bvh (trimesh.rs:1159-1171).
template <class T> void trimesh_build(TriMesh<T> &m) {
    std::vector<Aabb<T>> leaves(m.num_triangles());
    for (uint32_t i = 0; i < m.num_triangles(); ++i) {
        V3<T> a, b, c;
        m.tri(i, a, b, c);
        leaves[i] = triangle_aabb(a, b, c);
    }
    m.bvh = bvh_from_leaves(leaves);
}

struct RayCounters {
    uint64_t wide_loads = 0;   // iterations of the find_best loop (+ the partial root)
    uint64_t leaf_tests = 0;   // leaf_cost calls
    uint32_t max_stack = 0;
};

// RayCullingMode::check (ray_trimesh.rs:58-65) on the triangle's scaled
// normal (ab x ac, triangle.rs:324-328): culling 1 = IgnoreBackfaces keeps
// n.dir < 0, 2 = IgnoreFrontfaces keeps n.dir > 0, 0 keeps every triangle.
template <class T> inline bool culling_keeps(int culling, V3<T> a, V3<T> b, V3<T> c, V3<T> dir) {
    if (culling == 0) return true;
    T d = dot(cross(b - a, c - a), dir);
    return culling == 1 ? d < T(0) : d > T(0);
}

// TriMesh::cast_local_ray_and_get_normal (ray_trimesh.rs:16-35) through
// CompositeShapeRef (ray_composite_shape.rs:43-62) and Bvh::find_best
// (bvh_traverse.rs:335-417).  Returns false for None.  With culling != 0 it
// is TriMesh::cast_local_ray_with_culling (ray_trimesh.rs:152-178): the same
// traversal over TriMeshWithCulling, whose parts that fail the culling test
// are None (ray_trimesh.rs:97-127), and the same feature remap.
template <class T>
bool trimesh_cast_ray(const TriMesh<T> &m, const Ray<T> &ray, T max_toi, bool solid, RayHit<T> *out,
                      RayCounters *cnt, int culling = 0) {
    (void)solid;  // triangles ignore `solid`; Bvh nodes always use solid = true
    const auto &nodes = m.bvh.nodes;
    if (nodes.empty()) return false;
    auto leaf_cost = [&](uint32_t prim, T best, RayHit<T> *h) -> bool {
        if (cnt) cnt->leaf_tests++;
        V3<T> a, b, c;
        m.tri(prim, a, b, c);
        if (!culling_keeps(culling, a, b, c, ray.dir)) return false;
        return triangle_cast(a, b, c, ray, best, h);
    };
    uint32_t best_id = UINT32_MAX;
    RayHit<T> best_val{};
    bool have = false;
    if (nodes[0].right.leaf_count() == 0) {  // partial root, bvh_traverse.rs:348-359
        if (cnt) cnt->wide_loads++;
        const Node<T> &leaf = nodes[0].left;
        if (node_cast_ray(leaf, ray, max_toi) < max_toi) {
            RayHit<T> h;
            if (!leaf_cost(leaf.children, max_toi, &h)) return false;
            if (!(h.toi < max_toi)) return false;
            best_id = leaf.children;
            best_val = h;
            have = true;
        }
    } else {
        std::vector<uint32_t> stack;
        stack.reserve(32);
        T best_cost = max_toi;
        uint32_t curr = 0;
        for (;;) {
            if (cnt) cnt->wide_loads++;
            const Wide<T> &w = nodes[curr];
            const Node<T> *left = &w.left;
            const Node<T> *right = &w.right;
            T ls = node_cast_ray(*left, ray, best_cost);
            T rs = node_cast_ray(*right, ray, best_cost);
            if (ls > rs) {
                std::swap(ls, rs);
                std::swap(left, right);
            }
            bool found_next = false;
            if (ls < best_cost && ls != Consts<T>::max_value) {
                if (left->is_leaf()) {
                    RayHit<T> h;
                    if (leaf_cost(left->children, best_cost, &h) && h.toi < best_cost) {
                        best_val = h;
                        best_cost = h.toi;
                        best_id = left->children;
                        have = true;
                    }
                } else {
                    curr = left->children;
                    found_next = true;
                }
            }
            if (rs < best_cost && rs != Consts<T>::max_value) {
                if (right->is_leaf()) {
                    RayHit<T> h;
                    if (leaf_cost(right->children, best_cost, &h) && h.toi < best_cost) {
                        best_val = h;
                        best_cost = h.toi;
                        best_id = right->children;
                        have = true;
                    }
                } else if (found_next) {
                    stack.push_back(right->children);
                    if (cnt && stack.size() > cnt->max_stack) cnt->max_stack = (uint32_t)stack.size();
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
    }
    if (!have) return false;
    *out = best_val;
    // Feature remap, ray_trimesh.rs:25-33.
    out->fid = (best_val.fid == 1) ? best_id + m.num_triangles() : best_id;
    return true;
}

// Brute-force nearest hit over all triangles (test helper: pins find_best).
template <class T>
bool trimesh_cast_ray_brute(const TriMesh<T> &m, const Ray<T> &ray, T max_toi, RayHit<T> *out, int culling = 0) {
    bool have = false;
    T best = max_toi;
    uint32_t bid = 0;
    RayHit<T> bv{};
    for (uint32_t i = 0; i < m.num_triangles(); ++i) {
        V3<T> a, b, c;
        m.tri(i, a, b, c);
        RayHit<T> h;
        if (culling_keeps(culling, a, b, c, ray.dir) && triangle_cast(a, b, c, ray, best, &h) && h.toi < best) {
            best = h.toi;
            bv = h;
            bid = i;
            have = true;
        }
    }
    if (!have) return false;
    *out = bv;
    out->fid = (bv.fid == 1) ? bid + m.num_triangles() : bid;
    return true;
}

// ---------------------------------------------------------------------------
// Isometries (nalgebra Isometry3 = UnitQuaternion + Translation), restated from
// nalgebra 0.34's published implementation (not present in the reference tree,
// hence "parity unpinned" at the last ulp; see the file header):
//   q * v       : t = (q.v x v) * 2 ; r = (t * q.w + q.v x t) + v
//   iso * p     : q * p + t
//   q^-1        : conjugate (w, -x, -y, -z)
//   q1 * q2     : Hamilton product in nalgebra's term order
//   a.inv_mul(b): (q_a^-1 * (t_b - t_a), q_a^-1 * q_b)
// ---------------------------------------------------------------------------
template <class T> struct Quat {
    T x, y, z, w;  // nalgebra storage order [i, j, k, w]
};
template <class T> struct Iso {
    Quat<T> q;
    V3<T> t;
};
template <class T> inline V3<T> rotate(const Quat<T> &q, V3<T> v) {
    V3<T> qv{q.x, q.y, q.z};
    V3<T> t = cross(qv, v) * T(2);
    V3<T> c = cross(qv, t);
    return (t * q.w + c) + v;
}
template <class T> inline Quat<T> conj(const Quat<T> &q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> inline V3<T> inv_rotate(const Quat<T> &q, V3<T> v) { return rotate(conj(q), v); }
template <class T> inline V3<T> transform_point(const Iso<T> &m, V3<T> p) { return rotate(m.q, p) + m.t; }
template <class T> inline Quat<T> qmul(const Quat<T> &a, const Quat<T> &b) {
    Quat<T> r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return r;
}
template <class T> inline Iso<T> inv_mul(const Iso<T> &a, const Iso<T> &b) {
    Quat<T> ia = conj(a.q);
    return Iso<T>{qmul(ia, b.q), rotate(ia, b.t - a.t)};
}

// point_cloud_aabb (aabb_utils.rs:66-86), the AABB of a ConvexPolyhedron
// (aabb_convex_polyhedron.rs:8-10).
template <class T> inline Aabb<T> point_cloud_aabb(const Iso<T> &m, const V3<T> *pts, size_t n) {
    V3<T> w0 = transform_point(m, pts[0]);
    Aabb<T> r{w0, w0};
    for (size_t i = 1; i < n; ++i) {
        V3<T> w = transform_point(m, pts[i]);
        r.mins = vmin(r.mins, w);
        r.maxs = vmax(r.maxs, w);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Broadphase: leaf updates, refit and self-BVTT (bvh_insert.rs,
// bvh_refit.rs, bvh_traverse_bvtt.rs).
// ---------------------------------------------------------------------------
struct LeafSlot {
    uint32_t wide;
    bool right;
};

// leaf_node_indices of a Bvh: where each leaf lives.
template <class T> std::vector<LeafSlot> leaf_slots(const Bvh<T> &b, uint32_t n_leaves) {
    std::vector<LeafSlot> s(n_leaves, LeafSlot{0, false});
    for (uint32_t i = 0; i < b.nodes.size(); ++i) {
        const Wide<T> &w = b.nodes[i];
        if (w.left.leaf_count() == 1) s[w.left.children] = LeafSlot{i, false};
        if (w.right.leaf_count() == 1) s[w.right.children] = LeafSlot{i, true};
    }
    return s;
}

// insert_or_update_partially with change_detection_margin = 0 (bvh_insert.rs:209-231).
template <class T> void update_leaf(Bvh<T> &b, const LeafSlot &s, const Aabb<T> &a) {
    Node<T> &n = s.right ? b.nodes[s.wide].right : b.nodes[s.wide].left;
    n.mins = a.mins;
    n.maxs = a.maxs;
}

struct BvttStats {
    uint64_t wide_touched = 0;  // wide-node records read
};

template <class T> struct SelfBvtt {
    const Bvh<T> &b;
    std::vector<std::pair<uint32_t, uint32_t>> &out;
    BvttStats &st;
    std::vector<uint32_t> stack;

    void single_subtree(const Node<T> &node, uint32_t subtree) {  // bvh_traverse_bvtt.rs:145-204
        uint32_t curr = subtree;
        for (;;) {
            st.wide_touched++;
            const Wide<T> &w = b.nodes[curr];
            bool lc = node.intersects(w.left);
            bool rc = node.intersects(w.right);
            bool found = false;
            if (lc) {
                if (w.left.is_leaf()) out.push_back({node.children, w.left.children});
                else { curr = w.left.children; found = true; }
            }
            if (rc) {
                if (w.right.is_leaf()) out.push_back({node.children, w.right.children});
                else if (!found) { curr = w.right.children; found = true; }
                else stack.push_back(w.right.children);
            }
            if (!found) {
                if (stack.empty()) return;
                curr = stack.back();
                stack.pop_back();
            }
        }
    }
    void dispatch(bool check, const Node<T> &a, const Node<T> &c) {
        if (!check) return;
        if (a.is_leaf() && c.is_leaf()) out.push_back({a.children, c.children});
        else if (a.is_leaf()) single_subtree(a, c.children);
        else if (c.is_leaf()) single_subtree(c, a.children);
        else two_branches(a.children, c.children);
    }
    void two_branches(uint32_t x, uint32_t y) {  // bvh_traverse_bvtt.rs:89-142
        st.wide_touched += 2;
        const Wide<T> &n1 = b.nodes[x];
        const Wide<T> &n2 = b.nodes[y];
        bool ll = n1.left.intersects(n2.left);
        bool lr = n1.left.intersects(n2.right);
        bool rl = n1.right.intersects(n2.left);
        bool rr = n1.right.intersects(n2.right);
        dispatch(ll, n1.left, n2.left);
        dispatch(lr, n1.left, n2.right);
        dispatch(rl, n1.right, n2.left);
        dispatch(rr, n1.right, n2.right);
    }
    void self_node(uint32_t id) {  // bvh_traverse_bvtt.rs:38-87
        st.wide_touched++;
        const Wide<T> &n = b.nodes[id];
        bool lr = n.left.intersects(n.right);
        if (!n.left.is_leaf()) self_node(n.left.children);
        if (!n.right.is_leaf()) self_node(n.right.children);
        if (lr) {
            bool ll = n.left.is_leaf(), rl = n.right.is_leaf();
            if (ll && rl) out.push_back({n.left.children, n.right.children});
            else if (ll) single_subtree(n.left, n.right.children);
            else if (rl) single_subtree(n.right, n.left.children);
            else two_branches(n.left.children, n.right.children);
        }
    }
};

// traverse_bvtt_single_tree::<false> (bvh_traverse_bvtt.rs:19-31).
template <class T>
void self_pairs(const Bvh<T> &b, std::vector<std::pair<uint32_t, uint32_t>> &out, BvttStats &st) {
    if (b.nodes.empty() || b.nodes[0].right.leaf_count() == 0) return;
    SelfBvtt<T> v{b, out, st, {}};
    v.self_node(0);
}

}  // namespace oracle
