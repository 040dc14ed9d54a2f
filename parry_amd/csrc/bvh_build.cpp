// bvh_build.cpp -- host-side construction of parry's binned Bvh.
//
// The ray kernels must visit triangles in exactly parry's order to pick the
// same triangle on exact time-of-impact ties, so the tree topology has to be
// the one parry builds: Bvh::from_iter(BvhBuildStrategy::Binned, ..)
// (src/partitioning/bvh/bvh_tree.rs:1891-1955) = rebuild_range_binned
// (bvh_binned_build.rs:39-176) + refit's depth-first relayout
// (bvh_refit.rs:170-320).  This implementation is iterative: a LIFO task list
// produces depth-first preorder node ids directly (the relayout is then the
// identity), splits operate on a permutation of leaf ids with precomputed
// centres, and node boxes are filled bottom-up in one reverse sweep.
//
// Floating-point operations that decide the topology are kept in parry's
// order and precision (f32, no contraction): leaf centre (mins+maxs)*0.5,
// k1 = 8*(1-1e-5)/(hi-lo), bin = (k1*(c-k0)) as usize (saturating, NaN -> 0),
// cost = vol(L)*|L| + vol(R)*|R| with vol of an empty bin = (-inf)^3.
#include <cmath>
#include <cstring>
#include <limits>

#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBins = 8;
constexpr float kMaxF = 3.40282347e+38f;
constexpr uint32_t kChanged = 1u << 30;         // flag after refit (bvh_tree.rs:157, 192-198)
constexpr uint32_t kChangePending = 3u << 30;   // leaf flag at creation (bvh_tree.rs:158, 520-527)

struct Box {
    float lo[3], hi[3];
    static Box empty() { return Box{{kMaxF, kMaxF, kMaxF}, {-kMaxF, -kMaxF, -kMaxF}}; }
    void grow(const float *mn, const float *mx) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], mn[d]);
            hi[d] = std::fmax(hi[d], mx[d]);
        }
    }
    float volume() const {
        float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        return ex * ey * ez;
    }
};

inline uint64_t saturating_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}

struct Builder {
    const std::vector<LeafBox> &leaves;
    std::vector<float> cx, cy, cz;  // leaf centres
    std::vector<uint32_t> perm;
    std::vector<pq_bvh_node_wide> &out;

    const float *centre_axis(int axis) const { return axis == 0 ? cx.data() : (axis == 1 ? cy.data() : cz.data()); }

    // Splits perm[b, e) in place; returns the size of the left part.
    size_t split(size_t b, size_t e) {
        size_t len = e - b;
        float lo[3], hi[3];
        for (int d = 0; d < 3; ++d) lo[d] = hi[d] = centre_axis(d)[perm[b]];
        for (size_t i = b + 1; i < e; ++i) {
            uint32_t l = perm[i];
            float c[3] = {cx[l], cy[l], cz[l]};
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::fmin(lo[d], c[d]);
                hi[d] = std::fmax(hi[d], c[d]);
            }
        }
        // imax of the extents: first strict maximum.
        float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int axis = 0;
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        const float *cax = centre_axis(axis);
        const float k1 = (float)kBins * (1.0f - 1.0e-5f) / (hi[axis] - lo[axis]);
        const float k0 = lo[axis];
        auto bin_of = [&](uint32_t l) -> uint64_t { return saturating_usize(k1 * (cax[l] - k0)); };

        Box bin_box[kBins];
        uint32_t bin_cnt[kBins] = {0};
        for (int i = 0; i < kBins; ++i) bin_box[i] = Box::empty();
        for (size_t i = b; i < e; ++i) {
            uint32_t l = perm[i];
            uint64_t k = bin_of(l);
            if (k >= (uint64_t)kBins) k = kBins - 1;  // parry would panic here; unreachable for finite boxes
            bin_box[k].grow(leaves[l].mins, leaves[l].maxs);
            bin_cnt[k] += 1;
        }
        // Suffix unions: suffix[k] = bins k..7 (k >= 1).
        Box suf_box[kBins];
        uint32_t suf_cnt[kBins];
        suf_box[kBins - 1] = bin_box[kBins - 1];
        suf_cnt[kBins - 1] = bin_cnt[kBins - 1];
        for (int k = kBins - 2; k >= 1; --k) {
            suf_box[k] = suf_box[k + 1];
            suf_box[k].grow(bin_box[k].lo, bin_box[k].hi);
            suf_cnt[k] = suf_cnt[k + 1] + bin_cnt[k];
        }
        float best_cost = kMaxF;
        int best_plane = 0;
        uint32_t best_left = bin_cnt[0];
        Box pre_box = bin_box[0];
        uint32_t pre_cnt = bin_cnt[0];
        for (int p = 0; p < kBins - 1; ++p) {
            float cost = pre_box.volume() * (float)pre_cnt + suf_box[p + 1].volume() * (float)suf_cnt[p + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_plane = p;
                best_left = pre_cnt;
            }
            pre_box.grow(bin_box[p + 1].lo, bin_box[p + 1].hi);
            pre_cnt += bin_cnt[p + 1];
        }
        size_t mid = best_left;
        if (mid == 0 || mid == len) return len / 2;
        // Swap-based partition with the same visiting order as the reference:
        // a forward cursor over the left part finds a right-side leaf, a
        // forward cursor over the right part finds a left-side leaf, swap.
        size_t li = b, ri = b + mid;
        const size_t lend = b + mid;
        while (li != lend && ri != e) {
            while (li != lend && bin_of(perm[li]) <= (uint64_t)best_plane) ++li;
            if (li == lend) break;
            while (ri != e && bin_of(perm[ri]) > (uint64_t)best_plane) ++ri;
            if (ri == e) break;
            std::swap(perm[li], perm[ri]);
            ++li;
            ++ri;
        }
        return mid;
    }

    uint32_t run() {
        const uint32_t n = (uint32_t)leaves.size();
        struct Task {
            uint32_t parent;
            uint32_t right;  // 0: left slot, 1: right slot
            uint32_t b, e;
        };
        std::vector<Task> tasks;
        out.clear();
        out.reserve(n);
        out.push_back(pq_bvh_node_wide{});
        std::vector<uint32_t> level(1, 1);
        // The root's own range is split immediately; its children become tasks.
        {
            size_t mid = split(0, n);
            tasks.push_back(Task{0, 1, (uint32_t)mid, n});
            tasks.push_back(Task{0, 0, 0, (uint32_t)mid});
        }
        uint32_t depth = 1;
        while (!tasks.empty()) {
            Task t = tasks.back();
            tasks.pop_back();
            pq_bvh_node &slot = t.right ? out[t.parent].right : out[t.parent].left;
            if (t.e - t.b == 1) {
                uint32_t l = perm[t.b];
                std::memcpy(slot.mins, leaves[l].mins, 12);
                std::memcpy(slot.maxs, leaves[l].maxs, 12);
                slot.children = l;
                slot.data = 1u | kChanged;
                continue;
            }
            uint32_t id = (uint32_t)out.size();
            slot.children = id;
            out.push_back(pq_bvh_node_wide{});
            level.push_back(level[t.parent] + 1);
            if (level.back() > depth) depth = level.back();
            size_t mid = split(t.b, t.e);
            tasks.push_back(Task{id, 1, (uint32_t)(t.b + mid), t.e});
            tasks.push_back(Task{id, 0, t.b, (uint32_t)(t.b + mid)});
        }
        // Bottom-up boxes: in preorder a child's id is larger than its parent's,
        // so one reverse sweep sees every child before its parent.  Internal
        // slots still have data == 0 here; leaf slots were filled above.
        for (uint32_t i = (uint32_t)out.size(); i-- > 0;) {
            pq_bvh_node_wide &w = out[i];
            for (int s = 0; s < 2; ++s) {
                pq_bvh_node &c = s ? w.right : w.left;
                if (c.data != 0) continue;  // leaf, already filled
                const pq_bvh_node_wide &ch = out[c.children];
                for (int d = 0; d < 3; ++d) {
                    c.mins[d] = std::fmin(ch.left.mins[d], ch.right.mins[d]);
                    c.maxs[d] = std::fmax(ch.left.maxs[d], ch.right.maxs[d]);
                }
                c.data = (leaf_count(ch.left) + leaf_count(ch.right)) | kChanged;
            }
        }
        return depth;
    }
};

}  // namespace

void build_binned_bvh(const std::vector<LeafBox> &leaves, std::vector<pq_bvh_node_wide> &out, uint32_t *depth) {
    out.clear();
    const size_t n = leaves.size();
    auto leaf_node = [&](uint32_t l, uint32_t flags) {
        pq_bvh_node r;
        std::memcpy(r.mins, leaves[l].mins, 12);
        std::memcpy(r.maxs, leaves[l].maxs, 12);
        r.children = l;
        r.data = 1u | flags;
        return r;
    };
    if (n == 0) {
        *depth = 0;
        return;
    }
    if (n <= 2) {  // bvh_tree.rs:1914-1932: no rebuild, no refit (leaf flags stay pending)
        pq_bvh_node_wide w{};
        w.left = leaf_node(0, kChangePending);
        if (n == 2) w.right = leaf_node(1, kChangePending);
        out.push_back(w);
        *depth = 1;
        return;
    }
    Builder b{leaves, {}, {}, {}, {}, out};
    b.cx.resize(n);
    b.cy.resize(n);
    b.cz.resize(n);
    b.perm.resize(n);
    for (size_t i = 0; i < n; ++i) {
        b.cx[i] = (leaves[i].mins[0] + leaves[i].maxs[0]) * 0.5f;
        b.cy[i] = (leaves[i].mins[1] + leaves[i].maxs[1]) * 0.5f;
        b.cz[i] = (leaves[i].mins[2] + leaves[i].maxs[2]) * 0.5f;
        b.perm[i] = (uint32_t)i;
    }
    *depth = b.run();
}

uint32_t bvh_depth(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint32_t num_leaves, bool *ok) {
    *ok = false;
    if (num_nodes == 0) return 0;
    if (leaf_count(nodes[0].right) == 0) {
        *ok = num_nodes == 1 && is_leaf(nodes[0].left) && nodes[0].left.children < num_leaves;
        return 1;
    }
    struct Item {
        uint32_t id, level;
    };
    std::vector<Item> st{{0, 1}};
    std::vector<uint8_t> seen(num_nodes, 0);
    uint32_t depth = 0;
    uint64_t visited = 0;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        if (it.id >= num_nodes || seen[it.id]) return 0;
        seen[it.id] = 1;
        ++visited;
        if (it.level > depth) depth = it.level;
        const pq_bvh_node_wide &w = nodes[it.id];
        for (int s = 1; s >= 0; --s) {
            const pq_bvh_node &c = s ? w.right : w.left;
            if (is_leaf(c)) {
                if (c.children >= num_leaves) return 0;
            } else {
                st.push_back({c.children, it.level + 1});
            }
        }
    }
    *ok = visited == num_nodes;
    return depth;
}

}  // namespace pq
