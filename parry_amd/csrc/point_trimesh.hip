// point_trimesh.hip -- batched point projections onto a TriMesh through its Bvh (gfx950).
//
// TriMesh::project_local_point (src/query/point/point_composite_shape.rs:164-168
// -> CompositeShapeRef::project_local_point, :59-80): parry's best-first
// Bvh::find_best (bvh_traverse.rs:335-417) from max cost Real::MAX, with the
// point's distance to a child's AABB as node cost (Aabb::distance_to_local_point,
// solid = true, point_aabb.rs:135-146) and the distance to the point's
// projection on the triangle as leaf cost (Triangle::project_local_point,
// point_triangle.rs:54-291).  One lane per point, the ray kernel's node
// fetches (four 16-byte loads of the 64-byte wide record) and its LDS stack;
// the chosen triangle and the projection are bit-identical to parry3d's.
#include <cfloat>

#include "contact_gjk.h"
#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kPtBlock = 256;

// |max(mins - p, p - maxs, 0)| (nalgebra sup and norm).
__device__ __forceinline__ float aabb_point_dist(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                                 f3 p) {
    const float sx = fmaxf(fmaxf(mnx - p.x, p.x - mxx), 0.0f);
    const float sy = fmaxf(fmaxf(mny - p.y, p.y - mxy), 0.0f);
    const float sz = fmaxf(fmaxf(mnz - p.z, p.z - mxz), 0.0f);
    return sqrtf((sx * sx + sy * sy) + sz * sz);
}

struct PtBest {
    float cost;
    f3 p;
    bool inside;
    uint32_t leaf;
    bool have;
};

// A leaf of find_best: replaces the best only if strictly closer.
__device__ __forceinline__ void visit_leaf(const float4 *tri4, uint32_t leaf, f3 pt, bool solid, PtBest &b) {
    const float4 q0 = tri4[3 * leaf], q1 = tri4[3 * leaf + 1], q2 = tri4[3 * leaf + 2];
    bool inside = false;
    const f3 p = tri_project(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, f3{q1.z, q1.w, q2.x}, pt, solid, inside);
    const float cost = sqrtf(norm_sq3(sub3(pt, p)));  // na::distance
    if (cost < b.cost) {
        b.cost = cost;
        b.p = p;
        b.inside = inside;
        b.leaf = leaf;
        b.have = true;
    }
}

// COUNT = true also writes, per point, the wide-node records fetched and the
// triangles tested (the traffic model of the roofline figure).
template <bool COUNT>
__global__ __launch_bounds__(kPtBlock) void project_points_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                                  const TriRecord *__restrict__ tris,
                                                                  const float *__restrict__ points, uint64_t n,
                                                                  int solid_flag,
                                                                  pq_point_projection *__restrict__ out,
                                                                  uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t i = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t *stk = lds_stack + threadIdx.x;
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);
    const bool solid = solid_flag != 0;
    const f3 pt{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    PtBest b{FLT_MAX, f3{0.0f, 0.0f, 0.0f}, false, 0u, false};
    uint32_t n_wide = 0, n_tri = 0;

    const float4 root3 = node4[3];
    if ((__float_as_uint(root3.w) & 0x3fffffffu) == 0u) {  // partial root, bvh_traverse.rs:348-359
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        if (aabb_point_dist(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, pt) < FLT_MAX) {
            ++n_tri;
            visit_leaf(tri4, __float_as_uint(q0.w), pt, solid, b);
        }
    } else {
        uint32_t curr = 0, sp = 0;
        for (;;) {
            if (COUNT) ++n_wide;
            const float4 q0 = node4[4 * curr + 0];  // left mins, left children
            const float4 q1 = node4[4 * curr + 1];  // left maxs, left data
            const float4 q2 = node4[4 * curr + 2];  // right mins, right children
            const float4 q3 = node4[4 * curr + 3];  // right maxs, right data
            float ls = aabb_point_dist(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, pt);
            float rs = aabb_point_dist(q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, pt);
            uint32_t lc = __float_as_uint(q0.w), ld = __float_as_uint(q1.w);
            uint32_t rc = __float_as_uint(q2.w), rd = __float_as_uint(q3.w);
            if (ls > rs) {
                const float tf = ls;
                ls = rs;
                rs = tf;
                const uint32_t tc = lc;
                lc = rc;
                rc = tc;
                const uint32_t td = ld;
                ld = rd;
                rd = td;
            }
            bool found_next = false;
            if (ls < b.cost && ls != FLT_MAX) {
                if ((ld & 0x3fffffffu) == 1u) {
                    if (COUNT) ++n_tri;
                    visit_leaf(tri4, lc, pt, solid, b);
                } else {
                    curr = lc;
                    found_next = true;
                }
            }
            if (rs < b.cost && rs != FLT_MAX) {
                if ((rd & 0x3fffffffu) == 1u) {
                    if (COUNT) ++n_tri;
                    visit_leaf(tri4, rc, pt, solid, b);
                } else if (found_next) {
                    stk[sp * kPtBlock] = rc;
                    ++sp;
                } else {
                    curr = rc;
                    found_next = true;
                }
            }
            if (!found_next) {
                if (sp == 0) break;
                --sp;
                curr = stk[sp * kPtBlock];
            }
        }
    }
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }
    pq_point_projection o;
    if (b.have) {
        o.point[0] = b.p.x;
        o.point[1] = b.p.y;
        o.point[2] = b.p.z;
        o.is_inside = b.inside ? 1u : 0u;
        o.triangle = __float_as_uint(tri4[3 * b.leaf + 2].y);
    } else {
        o.point[0] = o.point[1] = o.point[2] = 0.0f;
        o.is_inside = 0u;
        o.triangle = PQ_NO_HIT;
    }
    out[i] = o;
}

}  // namespace

hipError_t launch_project_points(const DeviceMesh &m, const float *points, uint64_t n, int solid,
                                 pq_point_projection *out, uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    // The traversal stack never holds more than depth - 1 entries (as for rays).
    const size_t lds = (size_t)ray_stack_entries(m.depth) * kPtBlock * sizeof(uint32_t);
    const uint64_t blocks = (n + kPtBlock - 1) / kPtBlock;
    if (counts)
        hipLaunchKernelGGL(project_points_kernel<true>, dim3((uint32_t)blocks), dim3(kPtBlock), lds, stream, m.nodes,
                           m.tris, points, n, solid, out, counts);
    else
        hipLaunchKernelGGL(project_points_kernel<false>, dim3((uint32_t)blocks), dim3(kPtBlock), lds, stream,
                           m.nodes, m.tris, points, n, solid, out, counts);
    return hipGetLastError();
}

}  // namespace pq
