// ray_trimesh.hip -- batched ray casts against a TriMesh through its Bvh (gfx950).
//
// One lane per ray.  Each lane runs parry's best-first Bvh::find_best
// (src/partitioning/bvh/bvh_traverse.rs:335-417) with the ray/AABB slab test
// as node cost (bvh_tree.rs:1177-1181 -> ray_aabb.rs:12-49, solid = true) and
// the ray/triangle test as leaf cost (ray_triangle.rs:49-152), so the chosen
// triangle is bit-identical to parry3d's, ties included:
//   * children are scored against the current best, the cheaper one is taken
//     first, the left one on equal scores;
//   * a child is entered only if score < best && score != f32::MAX;
//   * a leaf replaces the best only if its toi is strictly smaller, and the
//     right child is scored against the best already updated by the left leaf.
// The node record is parry's 64-byte BvhNodeWide, fetched as four 16-byte
// loads; the short traversal stack (at most depth-1 entries) lives in LDS,
// strided by the block size so each lane's slot is bank-conflict free.
//
// Floating point: compiled with -ffp-contract=off; 1/x and sqrt are the
// correctly rounded IEEE operations, min/max are IEEE maxNum/minNum (the
// semantics of Rust's f32::min/max), so every value is the one parry computes.
#include <cfloat>

#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kRayBlock = 256;

struct LaneRay {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;  // 1/dir, the `denom` of ray_aabb.rs:22 hoisted out of the node loop
};

// Aabb::cast_local_ray with solid = true, mapped to BvhNode::cast_ray's
// unwrap_or(Real::MAX).  The reference exits as soon as tmin > tmax; because
// tmin only grows and tmax only shrinks, testing once after the three axes
// gives the same answer, so the three axes are evaluated branch-free.
__device__ __forceinline__ float slab_cost(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                           const LaneRay &r, float max_toi) {
    float tmin = 0.0f, tmax = max_toi;
    bool miss = false;
#define PQ_SLAB_AXIS(MN, MX, O, D, INV)                          \
    {                                                            \
        const bool z = (D == 0.0f);                              \
        miss = miss || (z && (O < MN || O > MX));                \
        float n = (MN - O) * INV;                                \
        float f = (MX - O) * INV;                                \
        const bool s = n > f;                                    \
        float lo = s ? f : n;                                    \
        float hi = s ? n : f;                                    \
        lo = z ? -INFINITY : lo;                                 \
        hi = z ? INFINITY : hi;                                  \
        tmin = fmaxf(tmin, lo);                                  \
        tmax = fminf(tmax, hi);                                  \
    }
    PQ_SLAB_AXIS(mnx, mxx, r.ox, r.dx, r.ix)
    PQ_SLAB_AXIS(mny, mxy, r.oy, r.dy, r.iy)
    PQ_SLAB_AXIS(mnz, mxz, r.oz, r.dz, r.iz)
#undef PQ_SLAB_AXIS
    return (miss || tmin > tmax) ? FLT_MAX : tmin;
}

struct TriHit {
    float toi;
    float nx, ny, nz;
    uint32_t back;  // FeatureId::Face(0) front / Face(1) back
};

// local_ray_intersection_with_triangle (ray_triangle.rs:70-152) followed by
// the `toi <= max_toi` acceptance of Triangle::cast_local_ray_and_get_normal.
__device__ __forceinline__ bool ray_triangle(const float4 q0, const float4 q1, const float4 q2, const LaneRay &r,
                                             float max_toi, TriHit &h) {
    const float ax = q0.x, ay = q0.y, az = q0.z;
    const float bx = q0.w, by = q1.x, bz = q1.y;
    const float cx = q1.z, cy = q1.w, cz = q2.x;
    const float abx = bx - ax, aby = by - ay, abz = bz - az;
    const float acx = cx - ax, acy = cy - ay, acz = cz - az;
    // n = ab x ac
    const float nx = aby * acz - abz * acy;
    const float ny = abz * acx - abx * acz;
    const float nz = abx * acy - aby * acx;
    float d = (nx * r.dx + ny * r.dy) + nz * r.dz;
    if (d == 0.0f) return false;
    const float apx = r.ox - ax, apy = r.oy - ay, apz = r.oz - az;
    const float t = (apx * nx + apy * ny) + apz * nz;
    if ((t < 0.0f && d < 0.0f) || (t > 0.0f && d > 0.0f)) return false;
    const uint32_t back = d < 0.0f ? 0u : 1u;
    d = fabsf(d);
    // e = -(dir x ap)
    const float ex = -(r.dy * apz - r.dz * apy);
    const float ey = -(r.dz * apx - r.dx * apz);
    const float ez = -(r.dx * apy - r.dy * apx);
    const float ac_e = (acx * ex + acy * ey) + acz * ez;
    const float ab_e = (abx * ex + aby * ey) + abz * ez;
    const bool neg = t < 0.0f;
    const float v = neg ? -ac_e : ac_e;
    if (v < 0.0f || v > d) return false;
    const float w = neg ? ab_e : -ab_e;
    if (w < 0.0f || v + w > d) return false;
    const float invd = 1.0f / d;
    const float toi = neg ? (-t * invd) : (t * invd);
    const float len = sqrtf((nx * nx + ny * ny) + nz * nz);
    float ux = nx / len, uy = ny / len, uz = nz / len;
    if (neg) {
        ux = -ux;
        uy = -uy;
        uz = -uz;
    }
    if (!(toi <= max_toi)) return false;
    h.toi = toi;
    h.nx = ux;
    h.ny = uy;
    h.nz = uz;
    h.back = back;
    return true;
}

__device__ __forceinline__ bool node_is_leaf(uint32_t data) { return (data & 0x3fffffffu) == 1u; }

// COUNT = true additionally writes, per ray, the number of wide-node records
// fetched and of triangles tested (the traffic model of the roofline figure).
template <bool COUNT>
__global__ __launch_bounds__(kRayBlock) void cast_rays_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                              const TriRecord *__restrict__ tris, uint32_t ntri,
                                                              const pq_ray *__restrict__ rays, uint64_t n,
                                                              float max_toi, pq_ray_hit *__restrict__ hits,
                                                              uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t i = (uint64_t)blockIdx.x * kRayBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t *stk = lds_stack + threadIdx.x;

    const float *rp = reinterpret_cast<const float *>(rays + i);
    const float2 r01 = *reinterpret_cast<const float2 *>(rp);
    const float2 r23 = *reinterpret_cast<const float2 *>(rp + 2);
    const float2 r45 = *reinterpret_cast<const float2 *>(rp + 4);
    LaneRay r;
    r.ox = r01.x;
    r.oy = r01.y;
    r.oz = r23.x;
    r.dx = r23.y;
    r.dy = r45.x;
    r.dz = r45.y;
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;

    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);

    bool have = false;
    TriHit best{};
    uint32_t best_leaf = 0;
    uint32_t n_wide = 0, n_tri = 0;

    const float4 root3 = node4[3];
    if ((__float_as_uint(root3.w) & 0x3fffffffu) == 0u) {
        // Single-leaf tree (bvh_traverse.rs:348-359).
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi) {
            const uint32_t leaf = __float_as_uint(q0.w);
            TriHit h;
            n_tri = 1;
            if (ray_triangle(tri4[3 * leaf], tri4[3 * leaf + 1], tri4[3 * leaf + 2], r, max_toi, h) &&
                h.toi < max_toi) {
                best = h;
                best_leaf = leaf;
                have = true;
            }
        }
    } else {
        float best_cost = max_toi;
        uint32_t curr = 0;
        uint32_t sp = 0;
        for (;;) {
            if (COUNT) ++n_wide;
            const float4 q0 = node4[4 * curr + 0];  // left mins, left children
            const float4 q1 = node4[4 * curr + 1];  // left maxs, left data
            const float4 q2 = node4[4 * curr + 2];  // right mins, right children
            const float4 q3 = node4[4 * curr + 3];  // right maxs, right data
            float ls = slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, best_cost);
            float rs = slab_cost(q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, r, best_cost);
            uint32_t lc = __float_as_uint(q0.w), ld = __float_as_uint(q1.w);
            uint32_t rc = __float_as_uint(q2.w), rd = __float_as_uint(q3.w);
            if (ls > rs) {
                float ts = ls;
                ls = rs;
                rs = ts;
                uint32_t tc = lc;
                lc = rc;
                rc = tc;
                uint32_t td = ld;
                ld = rd;
                rd = td;
            }
            bool found_next = false;
            if (ls < best_cost && ls != FLT_MAX) {
                if (node_is_leaf(ld)) {
                    TriHit h;
                    if (COUNT) ++n_tri;
                    if (ray_triangle(tri4[3 * lc], tri4[3 * lc + 1], tri4[3 * lc + 2], r, best_cost, h) &&
                        h.toi < best_cost) {
                        best = h;
                        best_cost = h.toi;
                        best_leaf = lc;
                        have = true;
                    }
                } else {
                    curr = lc;
                    found_next = true;
                }
            }
            if (rs < best_cost && rs != FLT_MAX) {
                if (node_is_leaf(rd)) {
                    TriHit h;
                    if (COUNT) ++n_tri;
                    if (ray_triangle(tri4[3 * rc], tri4[3 * rc + 1], tri4[3 * rc + 2], r, best_cost, h) &&
                        h.toi < best_cost) {
                        best = h;
                        best_cost = h.toi;
                        best_leaf = rc;
                        have = true;
                    }
                } else if (found_next) {
                    // sp stays below the LDS capacity: the host sizes it to
                    // depth - 1, the most entries a root-to-leaf path can leave
                    // pending (each pending entry sits one level deeper than the
                    // one below it).
                    stk[sp * kRayBlock] = rc;
                    ++sp;
                } else {
                    curr = rc;
                    found_next = true;
                }
            }
            if (!found_next) {
                if (sp == 0) break;
                --sp;
                curr = stk[sp * kRayBlock];
            }
        }
    }
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }

    pq_ray_hit out;
    if (have) {
        const uint32_t tri_id = __float_as_uint(tri4[3 * best_leaf + 2].y);
        out.time_of_impact = best.toi;
        out.normal[0] = best.nx;
        out.normal[1] = best.ny;
        out.normal[2] = best.nz;
        out.feature = best.back ? tri_id + ntri : tri_id;
    } else {
        out.time_of_impact = 0.0f;
        out.normal[0] = out.normal[1] = out.normal[2] = 0.0f;
        out.feature = PQ_NO_HIT;
    }
    hits[i] = out;
}

}  // namespace

hipError_t launch_cast_rays(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi, pq_ray_hit *hits,
                            uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t cap = m.depth > 1 ? m.depth - 1 : 1;
    const size_t lds = (size_t)cap * kRayBlock * sizeof(uint32_t);
    const uint64_t blocks = (n + kRayBlock - 1) / kRayBlock;
    if (counts)
        hipLaunchKernelGGL(cast_rays_kernel<true>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream, m.nodes,
                           m.tris, m.num_triangles, rays, n, max_toi, hits, counts);
    else
        hipLaunchKernelGGL(cast_rays_kernel<false>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream, m.nodes,
                           m.tris, m.num_triangles, rays, n, max_toi, hits, counts);
    return hipGetLastError();
}

}  // namespace pq
