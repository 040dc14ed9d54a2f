// ploc.hip -- Bvh::from_leaves(BvhBuildStrategy::Ploc, aabbs) built on the GPU.
//
// parry's build (bvh_tree.rs:1891-1955 -> rebuild_range_ploc,
// bvh_ploc_build.rs:10-93, then refit's depth-first relayout, bvh_refit.rs:170-320):
//   1. the AABB of the leaf centres ((mins + maxs) * 0.5) and its inverse
//      extents; each leaf's Morton key from its centre relative to that box,
//      widened to f64 and scaled by 2^21 (utils/morton.rs:13-40);
//   2. a stable sort of the leaves by key (sort_by_cached_key): an LSD radix
//      sort of (key, leaf index) pairs;
//   3. levels of merges until one node is left: every node i picks, among the
//      nodes at most 16 places away, the one whose merged box has the smallest
//      half area (the first on ties); mutual picks merge into a new wide node
//      (ids in order of the lower index, the last merge being the root, id 0)
//      that takes the lower index's place, and the other nodes carry over in
//      order.  One launch per step, with an exclusive scan giving each kept
//      node its place and each merge its id; the level's length stays on the
//      device (a one-lane kernel ends each level), and the host reads it back
//      every few levels to size the grids of the next ones;
//   4. the depth-first relayout: every wide node finds its position on its
//      path to the root (parent links recorded by the merges): a left child
//      sits right after its parent, a right child after the left subtree's
//      leaf count - 1 wide nodes; then one scatter writes every wide node at
//      its position with its children renamed and refit's flags (every slot
//      `changed`, bvh_tree.rs:192-198).
// Every float operation is parry's, in its order and precision, so the tree is
// the same bit for bit.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cfloat>
#include <vector>

#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBlock = 256;
constexpr int kRadius = 16;                  // SEARCH_RADIUS (bvh_ploc_build.rs:24)
constexpr uint32_t kPending = 3u << 30;      // BvhNodeData::with_leaf_count_and_pending_change
constexpr uint32_t kChanged = 1u << 30;      // after refit
constexpr uint32_t kCountMask = 0x3fffffffu;
constexpr uint32_t kReduceBlocks = 512;

struct Box6 {
    float lo[3], hi[3];
};

__device__ __forceinline__ uint64_t split_by_3(uint32_t a) {
    uint64_t x = (uint64_t)a & 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// Rust's `f64 as u32`: truncation, saturating, NaN -> 0.
__device__ __forceinline__ uint32_t as_u32(double f) {
    if (!(f > 0.0)) return 0u;
    if (f >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)f;
}

// Partial AABBs of the leaf centres; block b covers a strided share of the leaves.
__global__ __launch_bounds__(kBlock) void centre_box_partial_kernel(const float *__restrict__ aabbs, uint32_t n,
                                                                    Box6 *__restrict__ partial) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float *a = aabbs + 6 * (size_t)i;
        for (int d = 0; d < 3; ++d) {
            const float c = (a[d] + a[3 + d]) * 0.5f;
            lo[d] = fminf(lo[d], c);
            hi[d] = fmaxf(hi[d], c);
        }
    }
    __shared__ float s[6][kBlock];
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box6 b;
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = s[d][0];
            b.hi[d] = s[3 + d][0];
        }
        partial[blockIdx.x] = b;
    }
}

// Folds the partial boxes into partial[0] (one block; min and max are exact,
// so the order does not matter).
__global__ __launch_bounds__(kBlock) void centre_box_final_kernel(Box6 *__restrict__ partial, uint32_t nparts) {
    __shared__ float s[6][kBlock];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) {
        const Box6 q = partial[p];
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], q.lo[d]);
            hi[d] = fmaxf(hi[d], q.hi[d]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box6 b;
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = s[d][0];
            b.hi[d] = s[3 + d][0];
        }
        partial[0] = b;
    }
}

// Morton key of every leaf (bvh_ploc_build.rs:12-21) and the identity permutation.
__global__ __launch_bounds__(kBlock) void keys_kernel(const float *__restrict__ aabbs, uint32_t n,
                                                      const Box6 *__restrict__ cent,
                                                      unsigned long long *__restrict__ keys,
                                                      uint32_t *__restrict__ idx) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Box6 b = *cent;
    const float *a = aabbs + 6 * (size_t)i;
    uint64_t key = 0;
    for (int d = 0; d < 3; ++d) {
        const float inv = 1.0f / (b.hi[d] - b.lo[d]);
        const float c = ((a[d] + a[3 + d]) * 0.5f - b.lo[d]) * inv;
        key |= split_by_3(as_u32((double)c * 2097152.0)) << d;
    }
    keys[i] = key;
    idx[i] = i;
}

__global__ __launch_bounds__(kBlock) void init_leaves_kernel(const float *__restrict__ aabbs,
                                                             const uint32_t *__restrict__ order, uint32_t n,
                                                             pq_bvh_node *__restrict__ cur) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t l = order[j];
    const float *a = aabbs + 6 * (size_t)l;
    pq_bvh_node r;
    for (int d = 0; d < 3; ++d) {
        r.mins[d] = a[d];
        r.maxs[d] = a[3 + d];
    }
    r.children = l;
    r.data = 1u | kPending;
    cur[j] = r;
}

// Every node's merge candidate: the smallest half area of the merged box
// (Aabb::half_area, aabb.rs:473-476) within kRadius places, the first on ties.
__global__ __launch_bounds__(kBlock) void candidates_kernel(const pq_bvh_node *__restrict__ cur,
                                                            const uint32_t *__restrict__ level_len,
                                                            uint32_t *__restrict__ cand) {
    const uint32_t L = *level_len;
    if (L <= 1u || (uint64_t)blockIdx.x * kBlock >= L) return;  // whole blocks past the level (uniform exit)
    constexpr int kTile = kBlock + 2 * kRadius;
    __shared__ float lo[3][kTile], hi[3][kTile];
    const int64_t start = (int64_t)blockIdx.x * kBlock - kRadius;
    for (int t = threadIdx.x; t < kTile; t += kBlock) {
        const int64_t g = start + t;
        if (g >= 0 && g < (int64_t)L) {
            const pq_bvh_node nd = cur[g];
            for (int d = 0; d < 3; ++d) {
                lo[d][t] = nd.mins[d];
                hi[d][t] = nd.maxs[d];
            }
        }
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= L) return;
    const int ti = (int)threadIdx.x + kRadius;
    const uint32_t k0 = i >= (uint32_t)kRadius ? i - kRadius : 0u;
    const uint32_t k1 = i + kRadius < L - 1 ? i + kRadius : L - 1;
    float best = FLT_MAX;
    uint32_t bk = 0xffffffffu;
    for (uint32_t k = k0; k <= k1; ++k) {
        if (k == i) continue;
        const int tk = ti + (int)k - (int)i;
        const float ex = fmaxf(hi[0][ti], hi[0][tk]) - fminf(lo[0][ti], lo[0][tk]);
        const float ey = fmaxf(hi[1][ti], hi[1][tk]) - fminf(lo[1][ti], lo[1][tk]);
        const float ez = fmaxf(hi[2][ti], hi[2][tk]) - fminf(lo[2][ti], lo[2][tk]);
        const float sah = ex * (ey + ez) + ey * ez;
        if (sah < best) {
            best = sah;
            bk = k;
        }
    }
    cand[i] = bk;
}

// flags[i] = (kept << 32) | (merges here): i is kept unless it is the higher
// index of a mutual pair; it opens a merge if it is the lower one.
// Launched over bound + 1 entries (bound >= L): the entries from L on are zero.
__global__ __launch_bounds__(kBlock) void flags_kernel(const uint32_t *__restrict__ cand,
                                                       const uint32_t *__restrict__ level_len, uint32_t bound,
                                                       unsigned long long *__restrict__ flags) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i > bound) return;
    const uint32_t L = *level_len;
    if (L <= 1u) return;
    if (i >= L) {
        flags[i] = 0ull;
        return;
    }
    const uint32_t k = cand[i];
    const bool mutual = k < L && cand[k] == i;
    const bool keep = !mutual || i < k;
    const bool merge = mutual && i < k;
    flags[i] = ((unsigned long long)keep << 32) | (unsigned long long)merge;
}

__global__ __launch_bounds__(kBlock) void merge_kernel(const pq_bvh_node *__restrict__ cur,
                                                       const uint32_t *__restrict__ cand,
                                                       const unsigned long long *__restrict__ offs,
                                                       const uint32_t *__restrict__ level_len,
                                                       const uint32_t *__restrict__ next_id,
                                                       pq_bvh_node *__restrict__ next,
                                                       pq_bvh_node_wide *__restrict__ wide,
                                                       uint32_t *__restrict__ parent) {
    const uint32_t L = *level_len;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (L <= 1u || i >= L) return;
    const uint32_t base = *next_id;
    const uint32_t k = cand[i];
    const bool mutual = k < L && cand[k] == i;
    if (mutual && i > k) return;
    const unsigned long long o = offs[i];
    const uint32_t pos = (uint32_t)(o >> 32);
    if (!mutual) {
        next[pos] = cur[i];
        return;
    }
    const pq_bvh_node a = cur[i], b = cur[k];
    const uint32_t id = L == 2 ? 0u : base + (uint32_t)(o & 0xffffffffull);
    pq_bvh_node_wide w;
    w.left = a;
    w.right = b;
    wide[id] = w;
    // parent links of the wide children: parent wide id * 2 + side
    if ((a.data & kCountMask) > 1u) parent[a.children] = 2u * id;
    if ((b.data & kCountMask) > 1u) parent[b.children] = 2u * id + 1u;
    if (L == 2) return;
    pq_bvh_node m;  // BvhNode::merged (bvh_tree.rs:610-618, 200-204)
    for (int d = 0; d < 3; ++d) {
        m.mins[d] = fminf(a.mins[d], b.mins[d]);
        m.maxs[d] = fmaxf(a.maxs[d], b.maxs[d]);
    }
    m.children = id;
    m.data = ((a.data & kCountMask) + (b.data & kCountMask)) | (((a.data >> 30) | (b.data >> 30)) << 30);
    next[pos] = m;
}

// Ends a level: advances the next free wide id past the level's merges (the
// root, merged last, is id 0) and sets the next level's length.
// `state`: [0] level length, [1] next free wide id, [2] levels done, [3] error.
__global__ void finish_level_kernel(const unsigned long long *__restrict__ offs, uint32_t *__restrict__ state) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t L = state[0];
    if (L <= 1u) return;
    const unsigned long long tot = offs[L];
    const uint32_t kept = (uint32_t)(tot >> 32), merges = (uint32_t)(tot & 0xffffffffull);
    if (merges == 0u || kept >= L) {  // cannot happen: a mutual pair always exists
        state[3] = 1u;
        state[0] = 1u;
        return;
    }
    if (L != 2u) state[1] += merges;
    state[2] += 1u;
    state[0] = kept;
}

__global__ void init_state_kernel(uint32_t *state, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        state[0] = n;
        state[1] = 1u;
        state[2] = 0u;
        state[3] = 0u;
    }
}

// Depth-first position and level of every wide node, from the path to the root:
// a left child sits right after its parent, a right child after its parent and
// the left subtree's wide nodes (its leaf count - 1, none for a leaf).
__global__ __launch_bounds__(kBlock) void positions_kernel(const pq_bvh_node_wide *__restrict__ wide, uint32_t total,
                                                           const uint32_t *__restrict__ parent,
                                                           uint32_t *__restrict__ pos,
                                                           uint32_t *__restrict__ max_level) {
    const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= total) return;
    uint32_t p = 0, lv = 1, cur = w;
    while (cur != 0u && lv <= total) {  // at most total - 1 steps up
        const uint32_t link = parent[cur], par = link >> 1;
        uint32_t skip = 0;
        if (link & 1u) {
            const uint32_t lc = wide[par].left.data & kCountMask;
            skip = lc > 1u ? lc - 1u : 0u;
        }
        p += 1u + skip;
        ++lv;
        cur = par;
    }
    pos[w] = p;
    atomicMax(max_level, lv);
}

__device__ __forceinline__ pq_bvh_node relaid(pq_bvh_node c, const uint32_t *pos) {
    const uint32_t lc = c.data & kCountMask;
    if (lc > 1) c.children = pos[c.children];
    c.data = lc | kChanged;
    return c;
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(const pq_bvh_node_wide *__restrict__ wide, uint32_t total,
                                                         const uint32_t *__restrict__ pos,
                                                         pq_bvh_node_wide *__restrict__ out) {
    const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= total) return;
    const pq_bvh_node_wide nd = wide[w];
    pq_bvh_node_wide r;
    r.left = relaid(nd.left, pos);
    r.right = relaid(nd.right, pos);
    out[pos[w]] = r;
}

// One or two leaves: no rebuild, no refit (bvh_tree.rs:1914-1932).
__global__ void small_tree_kernel(const float *__restrict__ aabbs, uint32_t n, pq_bvh_node_wide *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    pq_bvh_node_wide w;
    for (int s = 0; s < 2; ++s) {
        pq_bvh_node &c = s ? w.right : w.left;
        if ((uint32_t)s < n) {
            for (int d = 0; d < 3; ++d) {
                c.mins[d] = aabbs[6 * s + d];
                c.maxs[d] = aabbs[6 * s + 3 + d];
            }
            c.children = (uint32_t)s;
            c.data = 1u | kPending;
        } else {
            for (int d = 0; d < 3; ++d) c.mins[d] = c.maxs[d] = 0.0f;
            c.children = 0u;
            c.data = 0u;
        }
    }
    out[0] = w;
}

uint32_t blocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

struct PlocWorkspace {
    Box6 *partial;
    unsigned long long *keys, *keys_sorted, *flags, *offs;
    uint32_t *idx, *idx_sorted, *cand, *pos, *parent, *max_level, *state;
    pq_bvh_node *cur, *next;
    pq_bvh_node_wide *wide;
    void *sort_temp, *scan_temp;
    size_t sort_bytes, scan_bytes, bytes;
};

PlocWorkspace carve(void *base, uint32_t n) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    char *ws = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = ws ? ws + off : nullptr;
        off += up(bytes);
        return p;
    };
    PlocWorkspace w{};
    w.partial = (Box6 *)take(kReduceBlocks * sizeof(Box6));
    w.keys = (unsigned long long *)take((size_t)n * 8);
    w.keys_sorted = (unsigned long long *)take((size_t)n * 8);
    w.flags = (unsigned long long *)take(((size_t)n + 1) * 8);
    w.offs = (unsigned long long *)take(((size_t)n + 1) * 8);
    w.idx = (uint32_t *)take((size_t)n * 4);
    w.idx_sorted = (uint32_t *)take((size_t)n * 4);
    w.cand = (uint32_t *)take((size_t)n * 4);
    w.pos = (uint32_t *)take((size_t)n * 4);
    w.parent = (uint32_t *)take((size_t)n * 4);
    w.max_level = (uint32_t *)take(4);
    w.state = (uint32_t *)take(4 * 4);
    w.cur = (pq_bvh_node *)take((size_t)n * sizeof(pq_bvh_node));
    w.next = (pq_bvh_node *)take((size_t)n * sizeof(pq_bvh_node));
    w.wide = (pq_bvh_node_wide *)take((size_t)n * sizeof(pq_bvh_node_wide));
    size_t sb = 0, cb = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sb, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                    (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 63);
    (void)rocprim::exclusive_scan(nullptr, cb, (unsigned long long *)nullptr, (unsigned long long *)nullptr, 0ull,
                                  (size_t)n + 1, rocprim::plus<unsigned long long>());
    w.sort_bytes = sb;
    w.scan_bytes = cb;
    w.sort_temp = take(sb);
    w.scan_temp = take(cb);
    w.bytes = off;
    return w;
}

}  // namespace

size_t ploc_workspace_bytes(uint32_t n) { return carve(nullptr, n < 3 ? 3 : n).bytes; }

hipError_t build_ploc_bvh(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t *depth, void *workspace,
                          hipStream_t stream) {
    *depth = 0;
    if (n == 0) return hipSuccess;
    if (n <= 2) {
        hipLaunchKernelGGL(small_tree_kernel, dim3(1), dim3(64), 0, stream, aabbs, n, out);
        *depth = 1;
        return hipGetLastError();
    }
    const PlocWorkspace w = carve(workspace, n);
    hipError_t e;
    const uint32_t parts = std::min<uint32_t>(kReduceBlocks, blocks(n));
    hipLaunchKernelGGL(centre_box_partial_kernel, dim3(parts), dim3(kBlock), 0, stream, aabbs, n, w.partial);
    hipLaunchKernelGGL(centre_box_final_kernel, dim3(1), dim3(kBlock), 0, stream, w.partial, parts);
    hipLaunchKernelGGL(keys_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream, aabbs, n, w.partial, w.keys, w.idx);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sb = w.sort_bytes;
    if ((e = rocprim::radix_sort_pairs(w.sort_temp, sb, w.keys, w.keys_sorted, w.idx, w.idx_sorted, (size_t)n, 0, 63,
                                       stream)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(init_leaves_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream, aabbs, w.idx_sorted, n, w.cur);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    // Merge levels, their length kept on the device: the host issues kLevelBatch
    // levels at a time with grids sized by the last length it read back, and the
    // kernels of a level past the end return at once.
    constexpr int kLevelBatch = 6;
    hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(64), 0, stream, w.state, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    pq_bvh_node *cur = w.cur, *next = w.next;
    uint32_t bound = n;
    uint32_t host_state[4] = {n, 1u, 0u, 0u};
    while (host_state[0] > 1u) {
        for (int k = 0; k < kLevelBatch; ++k) {
            hipLaunchKernelGGL(candidates_kernel, dim3(blocks(bound)), dim3(kBlock), 0, stream, cur, w.state, w.cand);
            hipLaunchKernelGGL(flags_kernel, dim3(blocks((uint64_t)bound + 1)), dim3(kBlock), 0, stream, w.cand,
                               w.state, bound, w.flags);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            size_t cb = w.scan_bytes;
            if ((e = rocprim::exclusive_scan(w.scan_temp, cb, w.flags, w.offs, 0ull, (size_t)bound + 1,
                                             rocprim::plus<unsigned long long>(), stream)) != hipSuccess)
                return e;
            hipLaunchKernelGGL(merge_kernel, dim3(blocks(bound)), dim3(kBlock), 0, stream, cur, w.cand, w.offs,
                               w.state, w.state + 1, next, w.wide, w.parent);
            hipLaunchKernelGGL(finish_level_kernel, dim3(1), dim3(64), 0, stream, w.offs, w.state);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            pq_bvh_node *t = cur;
            cur = next;
            next = t;
        }
        if ((e = hipMemcpyAsync(host_state, w.state, sizeof(host_state), hipMemcpyDeviceToHost, stream)) !=
            hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        bound = host_state[0];
    }
    if (host_state[3] != 0u || host_state[1] != n - 1) return hipErrorUnknown;
    if ((e = hipMemsetAsync(w.max_level, 0, sizeof(uint32_t), stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(positions_kernel, dim3(blocks(n - 1)), dim3(kBlock), 0, stream, w.wide, n - 1, w.parent, w.pos,
                       w.max_level);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(n - 1)), dim3(kBlock), 0, stream, w.wide, n - 1, w.pos, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(depth, w.max_level, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return e;
    return hipStreamSynchronize(stream);
}

}  // namespace pq
