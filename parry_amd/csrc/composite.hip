// composite.hip -- contacts between a TriMesh and convex shapes (gfx950).
//
// query::contact(pos1, TriMesh, pos2, g2) for a convex g2 takes the composite
// branch of DefaultQueryDispatcher::contact (default_query_dispatcher.rs:343-346)
// to CompositeShapeRef::contact_with_shape (contact_composite_shape_shape.rs:14-42):
// g2's AABB at pos12 = pos1^-1 pos2, loosened by the prediction, selects the
// mesh's triangles in Bvh::leaves order (bvh_traverse.rs:1-68, through
// Bvh::intersect_aabb, bvh_queries.rs:203-205); each triangle goes through the
// contact dispatch as a Triangle with that same pos12 (a TriMesh part has no pose
// of its own); the smallest dist wins, the first on ties; an Err or a None is
// skipped.
//
// Three steps on one stream:
//   1. one lane per shape computes its AABB in the mesh's frame and walks the
//      mesh's Bvh, counting, then (after an exclusive scan of the counts) writing
//      its (triangle, shape) candidate pairs in leaf order;
//   2. the batched contact (contact.hip) runs on every candidate, over a combined
//      shape table -- the mesh's triangles first, in the order of the Bvh's leaf
//      records, then the shapes -- in which every triangle is posed at the mesh's
//      pose, so each candidate's pos12 is computed exactly as parry's is;
//   3. one lane per shape keeps its first contact with the smallest dist.
#include <rocprim/device/device_scan.hpp>

#include "contact_gjk.h"
#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBlock = 256;

// The combined table's triangle records: one per Bvh leaf record (TriRecord),
// vertices a, b, c at 3t.
__global__ __launch_bounds__(kBlock) void mesh_shape_records_kernel(const TriRecord *__restrict__ tris, uint32_t ntri,
                                                                    ShapeRecord *__restrict__ rec,
                                                                    float4 *__restrict__ verts) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntri) return;
    const TriRecord r = tris[t];
    verts[3 * (size_t)t + 0] = make_float4(r.a[0], r.a[1], r.a[2], 0.0f);
    verts[3 * (size_t)t + 1] = make_float4(r.b[0], r.b[1], r.b[2], 0.0f);
    verts[3 * (size_t)t + 2] = make_float4(r.c[0], r.c[1], r.c[2], 0.0f);
    ShapeRecord s{};
    s.type = PQ_SHAPE_TRIANGLE;
    s.first_vertex = 3 * t;
    s.num_vertices = 3;
    rec[t] = s;
}

// Poses of the combined table: the mesh's pose for every triangle, then the shapes'.
__global__ __launch_bounds__(kBlock) void combined_poses_kernel(const float *__restrict__ mesh_pose,
                                                                const float *__restrict__ poses, uint32_t ntri,
                                                                uint32_t n, float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (uint64_t)ntri + n) return;
    const float *src = i < ntri ? mesh_pose : poses + 7 * (i - ntri);
    for (int k = 0; k < 7; ++k) out[7 * i + k] = src[k];
}

// BvhNode::intersects with a box (closed intervals, bvh_tree.rs:954-957).
__device__ __forceinline__ bool node_meets(float4 mn, float4 mx, f3 lo, f3 hi) {
    return (mn.x <= hi.x) & (mn.y <= hi.y) & (mn.z <= hi.z) & (mx.x >= lo.x) & (mx.y >= lo.y) & (mx.z >= lo.z);
}

// Step 1.  A node is named by (wide index << 1) | side; `stk` is this lane's
// LDS stack (stride kBlock), at most `depth` entries (one pending right sibling
// per level of the walk).
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void mesh_candidates_kernel(
    const pq_bvh_node_wide *__restrict__ nodes, const ShapeRecord *__restrict__ rec, const float4 *__restrict__ verts,
    const float *__restrict__ mesh_pose, const float *__restrict__ poses, uint32_t ntri, uint32_t n, float prediction,
    uint32_t *__restrict__ counts, const unsigned long long *__restrict__ offsets, uint32_t *__restrict__ pairs,
    uint64_t capacity) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    uint32_t *stk = lds_stack + threadIdx.x;
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const iso pos12 = iso_inv_mul(load_iso(mesh_pose), load_iso(poses + 7 * (size_t)q));
    f3 lo, hi;
    shape_aabb(rec[ntri + q], verts, pos12, lo, hi);
    lo = add3(lo, f3{-prediction, -prediction, -prediction});  // Aabb::loosened
    hi = add3(hi, f3{prediction, prediction, prediction});
    uint64_t out = WRITE ? offsets[q] : 0;
    uint32_t found = 0;
    uint32_t sp = 0;
    uint32_t next = 0;
    bool have = false;
    {
        const float4 l0 = node4[0], l1 = node4[1], r0 = node4[2], r1 = node4[3];
        if (node_meets(l0, l1, lo, hi)) {
            next = 0u;
            have = true;
        }
        if ((__float_as_uint(r1.w) & 0x3fffffffu) > 0u && node_meets(r0, r1, lo, hi)) stk[(sp++) * kBlock] = 1u;
    }
    for (;;) {
        // The pop as a select over a read every continuing lane makes (slot 0
        // exists: the host sizes the stack to depth + 1 entries), and the leaf
        // as an if/else: with a branch and a `continue` the compiler split the
        // loop into nested loops, and a wave's lanes that had to pop waited for
        // the others.
        if (!have && sp == 0) break;
        const uint32_t top = stk[(sp != 0u ? sp - 1u : 0u) * kBlock];
        next = have ? next : top;
        sp -= have ? 0u : 1u;
        have = false;
        const float4 mn = node4[2 * next], mx = node4[2 * next + 1];
        const uint32_t children = __float_as_uint(mn.w), data = __float_as_uint(mx.w);
        if ((data & 0x3fffffffu) == 1u) {  // a leaf: children is its TriRecord index
            if (WRITE) {
                if (out < capacity) {
                    pairs[2 * out] = children;
                    pairs[2 * out + 1] = ntri + q;
                }
                ++out;
            } else {
                ++found;
            }
        } else {
            const float4 cl0 = node4[4 * children], cl1 = node4[4 * children + 1];
            const float4 cr0 = node4[4 * children + 2], cr1 = node4[4 * children + 3];
            if (node_meets(cl0, cl1, lo, hi)) {
                next = 2u * children;
                have = true;
            }
            if (node_meets(cr0, cr1, lo, hi)) {
                if (!have) {
                    next = 2u * children + 1u;
                    have = true;
                } else {
                    stk[(sp++) * kBlock] = 2u * children + 1u;
                }
            }
        }
    }
    if (!WRITE) counts[q] = found;
}

__global__ void set_last_zero_kernel(uint32_t *counts, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[n] = 0u;
}
__global__ void copy_count_kernel(const unsigned long long *offsets, uint32_t n, unsigned long long *count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = offsets[n];
}

// Step 3.  A shape with a candidate beyond the capacity gets PQ_CONTACT_TRUNCATED.
__global__ __launch_bounds__(kBlock) void mesh_contact_select_kernel(
    const unsigned long long *__restrict__ offsets, uint32_t n, const uint32_t *__restrict__ pairs,
    const pq_contact *__restrict__ cand, const TriRecord *__restrict__ tris, uint64_t capacity,
    pq_contact *__restrict__ out, uint32_t *__restrict__ triangles) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    const uint64_t b = offsets[q], e = offsets[q + 1];
    pq_contact best{};
    best.has_contact = 0u;
    uint32_t tri = PQ_NO_HIT;
    if (e > capacity && e > b) {
        best.has_contact = PQ_CONTACT_TRUNCATED;
    } else {
        for (uint64_t k = b; k < e; ++k) {
            const pq_contact c = cand[k];
            if (c.has_contact != 1u) continue;  // None, or Err(Unsupported): skipped
            if (best.has_contact != 1u || c.dist < best.dist) {
                best = c;
                tri = tris[pairs[2 * k]].tri_id;
            }
        }
    }
    out[q] = best;
    if (triangles) triangles[q] = tri;
}

struct MeshContactWorkspace {
    float *poses;
    uint32_t *counts;
    unsigned long long *offsets;
    void *scan_temp;
    size_t scan_bytes;
    uint32_t *pairs;
    pq_contact *cand;
    unsigned long long *total;
    void *contact_ws;
    size_t bytes;
};

MeshContactWorkspace carve(void *base, uint32_t ntri, uint32_t n, uint64_t cap) {
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    char *ws = (char *)base;
    MeshContactWorkspace w{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = ws ? ws + off : nullptr;
        off += up(bytes);
        return p;
    };
    w.poses = (float *)take(((size_t)ntri + n) * 7 * sizeof(float));
    w.counts = (uint32_t *)take(((size_t)n + 1) * sizeof(uint32_t));
    w.offsets = (unsigned long long *)take(((size_t)n + 1) * sizeof(unsigned long long));
    size_t scan = 0;
    (void)rocprim::exclusive_scan(nullptr, scan, (const uint32_t *)nullptr, (unsigned long long *)nullptr, 0ull,
                                  (size_t)n + 1, rocprim::plus<unsigned long long>());
    w.scan_bytes = scan;
    w.scan_temp = take(scan);
    w.pairs = (uint32_t *)take((cap ? cap : 1) * 2 * sizeof(uint32_t));
    w.cand = (pq_contact *)take((cap ? cap : 1) * sizeof(pq_contact));
    w.total = (unsigned long long *)take(sizeof(unsigned long long));
    w.contact_ws = take(contact_workspace_bytes(cap ? cap : 1));
    w.bytes = off;
    return w;
}

}  // namespace

hipError_t build_mesh_shape_records(const DeviceMesh &m, ShapeRecord *rec, float4 *verts, hipStream_t stream) {
    if (m.num_triangles == 0) return hipSuccess;
    hipLaunchKernelGGL(mesh_shape_records_kernel, dim3((m.num_triangles + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       stream, m.tris, m.num_triangles, rec, verts);
    return hipGetLastError();
}

size_t mesh_contact_workspace_bytes(uint32_t ntri, uint32_t n, uint64_t capacity) {
    return carve(nullptr, ntri, n, capacity).bytes;
}

hipError_t launch_mesh_contacts(const DeviceMesh &m, const DeviceShapes &combined, uint32_t n, const float *mesh_pose,
                                const float *poses, float prediction, uint64_t capacity, pq_contact *out,
                                uint32_t *triangles, unsigned long long *num_candidates, void *workspace,
                                hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t ntri = m.num_triangles;
    const MeshContactWorkspace w = carve(workspace, ntri, n, capacity);
    const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    const size_t lds = ((size_t)m.depth + 1) * kBlock * sizeof(uint32_t);
    hipLaunchKernelGGL(combined_poses_kernel, dim3((uint32_t)(((uint64_t)ntri + n + kBlock - 1) / kBlock)), block, 0,
                       stream, mesh_pose, poses, ntri, n, w.poses);
    hipLaunchKernelGGL(mesh_candidates_kernel<false>, grid, block, lds, stream, m.nodes, combined.rec, combined.verts,
                       mesh_pose, poses, ntri, n, prediction, w.counts, nullptr, nullptr, 0ull);
    hipLaunchKernelGGL(set_last_zero_kernel, dim3(1), dim3(64), 0, stream, w.counts, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = w.scan_bytes;
    e = rocprim::exclusive_scan(w.scan_temp, bytes, w.counts, w.offsets, 0ull, (size_t)n + 1,
                                rocprim::plus<unsigned long long>(), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mesh_candidates_kernel<true>, grid, block, lds, stream, m.nodes, combined.rec, combined.verts,
                       mesh_pose, poses, ntri, n, prediction, nullptr, w.offsets, w.pairs, capacity);
    hipLaunchKernelGGL(copy_count_kernel, dim3(1), dim3(64), 0, stream, w.offsets, n, w.total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (capacity) {
        e = launch_contacts(combined, w.poses, w.pairs, capacity, (const uint64_t *)w.total, prediction, nullptr, w.cand, nullptr,
                            w.contact_ws, stream, kContactShortEpa);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(mesh_contact_select_kernel, grid, block, 0, stream, w.offsets, n, w.pairs, w.cand, m.tris,
                       capacity, out, triangles);
    if (num_candidates)
        hipLaunchKernelGGL(copy_count_kernel, dim3(1), dim3(64), 0, stream, w.offsets, n, num_candidates);
    return hipGetLastError();
}

}  // namespace pq
