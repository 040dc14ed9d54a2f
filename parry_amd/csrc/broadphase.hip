// broadphase.hip -- Bvh leaf update + refit and self-overlap pairs (gfx950).
//
// Refit (Bvh::insert_or_update_partially with margin 0 for every leaf, then
// Bvh::refit, src/partitioning/bvh/bvh_insert.rs:209-231, bvh_refit.rs:170-320):
// one launch writes every leaf's new box into its slot; then the wide nodes are
// refitted level by level, a level being the nodes of one height above the
// leaves (computed once, when the tree is built), so each node merges two
// child records that earlier levels finished -- min/max, leaf counts, change
// flags, exactly BvhNode::merged (bvh_tree.rs:610-618).  Wide levels get one
// launch each; the narrow top levels run in one workgroup with a barrier
// between levels.  Topology never changes, and the stored layout is already
// depth-first, so parry's relayout is the identity.
//
// Pairs (Bvh::traverse_bvtt_single_tree::<false>, bvh_traverse_bvtt.rs:19-204):
// parry reports every pair of leaves whose boxes overlap, which is the set
// {(i, j) : i < j, intersects(i, j)} because every ancestor box contains its
// descendants exactly.  One lane per leaf i walks the tree from the root with
// an LDS stack and reports the leaves j > i it overlaps.  The list is written
// without atomics: a counting pass, an exclusive scan of the per-leaf counts
// (rocPRIM), and a writing pass that places leaf i's pairs at its offset.
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kFlagChanged = 1u, kFlagPending = 3u;

// Leaf update (insert_or_update_partially with margin 0, bvh_insert.rs:209-231):
// the leaf's record takes its new box and resolves a pending change flag
// (bvh_tree.rs:192-198).
__global__ __launch_bounds__(kBlock) void refit_leaves_kernel(pq_bvh_node_wide *__restrict__ nodes,
                                                              const uint32_t *__restrict__ leaf_slot,
                                                              const float *__restrict__ aabbs, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = leaf_slot[i];
    const uint32_t wide = slot >> 1;
    pq_bvh_node *leaf = (slot & 1u) ? &nodes[wide].right : &nodes[wide].left;
    const float *a = aabbs + 6 * (size_t)i;
    leaf->mins[0] = a[0];
    leaf->mins[1] = a[1];
    leaf->mins[2] = a[2];
    leaf->maxs[0] = a[3];
    leaf->maxs[1] = a[4];
    leaf->maxs[2] = a[5];
    const uint32_t d = leaf->data;
    leaf->data = (d & 0x3fffffffu) | (((d >> 30) == kFlagPending ? kFlagChanged : 0u) << 30);
}

// The record of wide node w in its parent: the union of w's two child records
// (bvh_refit.rs:170-320; min/max are exact, so any order gives parry's boxes),
// leaf counts summed, change flags or-ed.
__device__ __forceinline__ void refit_node(pq_bvh_node_wide *nodes, const uint32_t *parent_slot, uint32_t w) {
    const pq_bvh_node l = nodes[w].left, r = nodes[w].right;
    const uint32_t ps = parent_slot[w];
    pq_bvh_node *dst = (ps & 1u) ? &nodes[ps >> 1].right : &nodes[ps >> 1].left;
    for (int k = 0; k < 3; ++k) {
        dst->mins[k] = fminf(l.mins[k], r.mins[k]);
        dst->maxs[k] = fmaxf(l.maxs[k], r.maxs[k]);
    }
    dst->data = ((l.data & 0x3fffffffu) + (r.data & 0x3fffffffu)) | (((l.data >> 30) | (r.data >> 30)) << 30);
}

// One level of the refit: the nodes of one height, in parallel.
__global__ __launch_bounds__(kBlock) void refit_level_kernel(pq_bvh_node_wide *__restrict__ nodes,
                                                             const uint32_t *__restrict__ parent_slot,
                                                             const uint32_t *__restrict__ ids, uint32_t count) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k < count) refit_node(nodes, parent_slot, ids[k]);
}

// The narrow top levels in one workgroup, a barrier between levels.  The
// workgroup runs on one CU, whose waves share its L1, so the barrier's
// workgroup-scope release/acquire is enough for one level to see the records
// the previous level wrote; nothing has to be written back to or invalidated
// in the L2 between levels.
constexpr uint32_t kTopBlock = 1024;
constexpr uint32_t kTopLevelMax = 4096;  // levels at most this wide go to the one-workgroup kernel
__global__ __launch_bounds__(kTopBlock) void refit_top_kernel(pq_bvh_node_wide *nodes,
                                                              const uint32_t *__restrict__ parent_slot,
                                                              const uint32_t *__restrict__ ids,
                                                              const uint32_t *__restrict__ level_offsets,
                                                              uint32_t first_level, uint32_t last_level) {
    for (uint32_t h = first_level; h <= last_level; ++h) {
        const uint32_t begin = level_offsets[h - 1], end = level_offsets[h];
        for (uint32_t k = begin + threadIdx.x; k < end; k += kTopBlock) refit_node(nodes, parent_slot, ids[k]);
        __syncthreads();
    }
}

__device__ __forceinline__ bool overlaps(const float q[6], float4 mn, float4 mx) {
    // BvhNode::intersects: mins <= other.maxs && maxs >= other.mins on every axis.
    return (q[0] <= mx.x) & (q[1] <= mx.y) & (q[2] <= mx.z) & (q[3] >= mn.x) & (q[4] >= mn.y) & (q[5] >= mn.z);
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void self_pairs_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                            const uint32_t *__restrict__ leaf_slot, uint32_t n,
                                                            uint32_t *__restrict__ counts,
                                                            const unsigned long long *__restrict__ offsets,
                                                            uint32_t *__restrict__ pairs, uint64_t capacity) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t *stk = lds_stack + threadIdx.x;
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const uint32_t slot = leaf_slot[i];
    const float4 a0 = node4[4 * (slot >> 1) + 2 * (slot & 1u)];
    const float4 a1 = node4[4 * (slot >> 1) + 2 * (slot & 1u) + 1];
    const float q[6] = {a0.x, a0.y, a0.z, a1.x, a1.y, a1.z};
    uint64_t base = 0;
    if (WRITE) base = offsets[i];
    uint32_t found = 0;
    uint32_t curr = 0, sp = 0;
    for (;;) {
        const float4 l0 = node4[4 * curr + 0], l1 = node4[4 * curr + 1];
        const float4 r0 = node4[4 * curr + 2], r1 = node4[4 * curr + 3];
        const uint32_t lc = __float_as_uint(l0.w), ld = __float_as_uint(l1.w);
        const uint32_t rc = __float_as_uint(r0.w), rd = __float_as_uint(r1.w);
        const bool lhit = overlaps(q, l0, l1);
        const bool rhit = overlaps(q, r0, r1) & ((rd & 0x3fffffffu) != 0u);
        bool next = false;
        if (lhit) {
            if ((ld & 0x3fffffffu) == 1u) {
                if (lc > i) {
                    if (WRITE && base + found < capacity) {
                        pairs[2 * (base + found)] = i;
                        pairs[2 * (base + found) + 1] = lc;
                    }
                    ++found;
                }
            } else {
                curr = lc;
                next = true;
            }
        }
        if (rhit) {
            if ((rd & 0x3fffffffu) == 1u) {
                if (rc > i) {
                    if (WRITE && base + found < capacity) {
                        pairs[2 * (base + found)] = i;
                        pairs[2 * (base + found) + 1] = rc;
                    }
                    ++found;
                }
            } else if (next) {
                // at most depth-1 pending entries (each one level deeper than the
                // one below it); the host sizes the LDS stack to that
                stk[sp * kBlock] = rc;
                ++sp;
            } else {
                curr = rc;
                next = true;
            }
        }
        // The pop as a select over a read every continuing lane makes (slot 0
        // exists: the host sizes the stack to at least one entry).  Written as
        // a branch, the compiler split the loop in two with the pop outside the
        // inner loop, so a wave's lanes that had to pop waited for the others.
        if (!next && sp == 0) break;
        const uint32_t top = stk[(sp != 0u ? sp - 1u : 0u) * kBlock];
        curr = next ? curr : top;
        sp -= next ? 0u : 1u;
    }
    if (!WRITE) counts[i] = found;
}

__global__ void copy_total_kernel(const unsigned long long *offsets, uint32_t n, unsigned long long *count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = offsets[n];
}

}  // namespace

hipError_t launch_bvh_refit(const DeviceBvh &b, const float *aabbs, hipStream_t stream) {
    if (b.num_leaves == 0) return hipSuccess;
    hipLaunchKernelGGL(refit_leaves_kernel, dim3((b.num_leaves + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                       b.nodes, b.leaf_slot, aabbs, b.num_leaves);
    hipError_t e = hipGetLastError();
    uint32_t h = 1;
    for (; e == hipSuccess && h <= b.num_levels; ++h) {
        const uint32_t begin = b.level_offsets_host[h - 1], count = b.level_offsets_host[h] - begin;
        if (count <= kTopLevelMax) break;
        hipLaunchKernelGGL(refit_level_kernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, b.nodes,
                           b.parent_slot, b.level_nodes + begin, count);
        e = hipGetLastError();
    }
    if (e == hipSuccess && h <= b.num_levels) {
        hipLaunchKernelGGL(refit_top_kernel, dim3(1), dim3(kTopBlock), 0, stream, b.nodes, b.parent_slot,
                           b.level_nodes, b.level_offsets, h, b.num_levels);
        e = hipGetLastError();
    }
    return e;
}

hipError_t bvh_scan_temp_bytes(uint32_t num_leaves, size_t *bytes) {
    return rocprim::exclusive_scan(nullptr, *bytes, (const uint32_t *)nullptr, (unsigned long long *)nullptr,
                                   0ull, (size_t)num_leaves + 1, rocprim::plus<unsigned long long>());
}

hipError_t launch_bvh_self_pairs(const DeviceBvh &b, uint32_t *pairs, uint64_t capacity, unsigned long long *count,
                                 hipStream_t stream) {
    const uint32_t n = b.num_leaves;
    if (n < 2) {
        // traverse_bvtt_single_tree returns at once (bvh_traverse_bvtt.rs:24-27).
        hipLaunchKernelGGL(copy_total_kernel, dim3(1), dim3(64), 0, stream, b.pair_offsets, 0u, count);
        return hipGetLastError();
    }
    const uint32_t cap = b.depth > 1 ? b.depth - 1 : 1;
    const size_t lds = (size_t)cap * kBlock * sizeof(uint32_t);
    const dim3 grid((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(self_pairs_kernel<false>, grid, dim3(kBlock), lds, stream, b.nodes, b.leaf_slot, n,
                       b.pair_counts, b.pair_offsets, pairs, capacity);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = b.scan_temp_bytes;
    e = rocprim::exclusive_scan(b.scan_temp, bytes, b.pair_counts, b.pair_offsets, 0ull, (size_t)n + 1,
                                rocprim::plus<unsigned long long>(), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(self_pairs_kernel<true>, grid, dim3(kBlock), lds, stream, b.nodes, b.leaf_slot, n,
                       b.pair_counts, b.pair_offsets, pairs, capacity);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(copy_total_kernel, dim3(1), dim3(64), 0, stream, b.pair_offsets, n, count);
    return hipGetLastError();
}

}  // namespace pq
