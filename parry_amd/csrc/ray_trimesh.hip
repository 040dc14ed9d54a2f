// ray_trimesh.hip -- batched ray casts against a TriMesh through its Bvh (gfx950).
//
// One lane per ray at a time.  Each lane runs parry's best-first Bvh::find_best
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
// Two lane mappings (chosen per mesh by the tuner, pq_trimesh_set_tuning):
//   * static: one launch lane per ray; a wave lasts as long as its slowest ray;
//   * persistent: a grid sized to the resident waves; each wave pulls chunks
//     of ray indices from a queue in device memory, and every iteration the
//     lanes whose ray has finished (found with a wave ballot) take the next
//     rays of the chunk, ranked by a masked popcount, so the wave stays full
//     until the queue drains.  The queue counter resets itself: the last
//     wave to finish zeroes it for the next launch on the stream.
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
    // Bitwise (not short-circuit) boolean algebra keeps the test free of
    // exec-mask branches: every operand is a plain compare.
    unsigned miss = 0u;
#define PQ_SLAB_AXIS(MN, MX, O, D, INV)                               \
    {                                                                 \
        const unsigned z = (unsigned)(D == 0.0f);                      \
        miss |= z & ((unsigned)(O < MN) | (unsigned)(O > MX));         \
        const float n = (MN - O) * INV;                                \
        const float f = (MX - O) * INV;                                \
        const bool s = n > f;                                          \
        float lo = s ? f : n;                                          \
        float hi = s ? n : f;                                          \
        lo = z ? -INFINITY : lo;                                       \
        hi = z ? INFINITY : hi;                                        \
        tmin = fmaxf(tmin, lo);                                        \
        tmax = fminf(tmax, hi);                                        \
    }
    PQ_SLAB_AXIS(mnx, mxx, r.ox, r.dx, r.ix)
    PQ_SLAB_AXIS(mny, mxy, r.oy, r.dy, r.iy)
    PQ_SLAB_AXIS(mnz, mxz, r.oz, r.dz, r.iz)
#undef PQ_SLAB_AXIS
    miss |= (unsigned)(tmin > tmax);
    return miss ? FLT_MAX : tmin;
}

struct TriHit {
    float toi;
    float nx, ny, nz;
    uint32_t back;  // FeatureId::Face(0) front / Face(1) back
};

// local_ray_intersection_with_triangle (ray_triangle.rs:70-152) followed by
// the `toi <= max_toi` acceptance of Triangle::cast_local_ray_and_get_normal.
// CULL 1 / 2 first drops the triangles RayCullingMode::IgnoreBackfaces /
// IgnoreFrontfaces skips (ray_trimesh.rs:58-65): its test is the sign of the
// scaled normal (ab x ac) dotted with the ray direction, which is `d` below.
template <int CULL = 0>
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
    if (CULL == 1 && !(d < 0.0f)) return false;
    if (CULL == 2 && !(d > 0.0f)) return false;
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

struct Traversal {
    float best_cost;
    TriHit best;
    uint32_t best_leaf;
    uint32_t curr;
    uint32_t sp;
    bool have;
};

__device__ __forceinline__ LaneRay load_ray(const pq_ray *rays, uint64_t i) {
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
    return r;
}

__device__ __forceinline__ void start_traversal(Traversal &t, float max_toi) {
    t.best_cost = max_toi;
    t.best = TriHit{};
    t.best_leaf = 0;
    t.curr = 0;
    t.sp = 0;
    t.have = false;
}

// Tests one leaf (a triangle record) against the current best.
template <bool COUNT, int CULL = 0>
__device__ __forceinline__ void visit_leaf(const float4 *tri4, uint32_t leaf, const LaneRay &r, Traversal &t,
                                           uint32_t &n_tri) {
    TriHit h;
    if (COUNT) ++n_tri;
    if (ray_triangle<CULL>(tri4[3 * leaf], tri4[3 * leaf + 1], tri4[3 * leaf + 2], r, t.best_cost, h) &&
        h.toi < t.best_cost) {
        t.best = h;
        t.best_cost = h.toi;
        t.best_leaf = leaf;
        t.have = true;
    }
}

// One iteration of find_best's loop (bvh_traverse.rs:361-416) on wide node
// t.curr.  Returns true when the traversal has finished.  `stk` is this lane's
// LDS stack (stride kRayBlock); it never holds more than depth-1 entries, the
// most a root-to-leaf path can leave pending (each pending entry sits one
// level deeper than the one below it), and the host sizes it to that.
template <bool COUNT, int CULL = 0>
__device__ __forceinline__ bool find_best_step(const float4 *node4, const float4 *tri4, const LaneRay &r,
                                               Traversal &t, uint32_t *stk, uint32_t &n_wide, uint32_t &n_tri) {
    if (COUNT) ++n_wide;
    const float4 q0 = node4[4 * t.curr + 0];  // left mins, left children
    const float4 q1 = node4[4 * t.curr + 1];  // left maxs, left data
    const float4 q2 = node4[4 * t.curr + 2];  // right mins, right children
    const float4 q3 = node4[4 * t.curr + 3];  // right maxs, right data
    float ls = slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, t.best_cost);
    float rs = slab_cost(q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, r, t.best_cost);
    uint32_t lc = __float_as_uint(q0.w), ld = __float_as_uint(q1.w);
    uint32_t rc = __float_as_uint(q2.w), rd = __float_as_uint(q3.w);
    const bool swap = ls > rs;
    {
        const float a = swap ? rs : ls, b = swap ? ls : rs;
        ls = a;
        rs = b;
        const uint32_t c0 = swap ? rc : lc, c1 = swap ? lc : rc;
        lc = c0;
        rc = c1;
        const uint32_t d0 = swap ? rd : ld, d1 = swap ? ld : rd;
        ld = d0;
        rd = d1;
    }
    bool found_next = false;
    if (ls < t.best_cost && ls != FLT_MAX) {
        if (node_is_leaf(ld)) {
            visit_leaf<COUNT, CULL>(tri4, lc, r, t, n_tri);
        } else {
            t.curr = lc;
            found_next = true;
        }
    }
    if (rs < t.best_cost && rs != FLT_MAX) {
        if (node_is_leaf(rd)) {
            visit_leaf<COUNT, CULL>(tri4, rc, r, t, n_tri);
        } else if (found_next) {
            stk[t.sp * kRayBlock] = rc;
            ++t.sp;
        } else {
            t.curr = rc;
            found_next = true;
        }
    }
    if (!found_next) {
        if (t.sp == 0) return true;
        --t.sp;
        t.curr = stk[t.sp * kRayBlock];
    }
    return false;
}

__device__ __forceinline__ void write_hit(pq_ray_hit *hits, uint64_t i, const Traversal &t, const float4 *tri4,
                                          uint32_t ntri) {
    pq_ray_hit out;
    if (t.have) {
        const uint32_t tri_id = __float_as_uint(tri4[3 * t.best_leaf + 2].y);
        out.time_of_impact = t.best.toi;
        out.normal[0] = t.best.nx;
        out.normal[1] = t.best.ny;
        out.normal[2] = t.best.nz;
        out.feature = t.best.back ? tri_id + ntri : tri_id;  // ray_trimesh.rs:25-33
    } else {
        out.time_of_impact = 0.0f;
        out.normal[0] = out.normal[1] = out.normal[2] = 0.0f;
        out.feature = PQ_NO_HIT;
    }
    hits[i] = out;
}

// Static mapping; also serves trees whose root holds a single leaf
// (bvh_traverse.rs:348-359).  COUNT = true additionally writes, per ray, the
// number of wide-node records fetched and of triangles tested (the traffic
// model of the roofline figure).  CULL selects the culling variant
// (TriMesh::cast_local_ray_with_culling, ray_trimesh.rs:152-178).
template <bool COUNT, int CULL = 0>
__global__ __launch_bounds__(kRayBlock) void cast_rays_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                              const TriRecord *__restrict__ tris, uint32_t ntri,
                                                              const pq_ray *__restrict__ rays, uint64_t n,
                                                              float max_toi, pq_ray_hit *__restrict__ hits,
                                                              uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t i = (uint64_t)blockIdx.x * kRayBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t *stk = lds_stack + threadIdx.x;
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);
    const LaneRay r = load_ray(rays, i);
    Traversal t;
    start_traversal(t, max_toi);
    uint32_t n_wide = 0, n_tri = 0;

    const float4 root3 = node4[3];
    if ((__float_as_uint(root3.w) & 0x3fffffffu) == 0u) {
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
            visit_leaf<true, CULL>(tri4, __float_as_uint(q0.w), r, t, n_tri);
    } else {
        while (!find_best_step<COUNT, CULL>(node4, tri4, r, t, stk, n_wide, n_tri)) {
        }
    }
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }
    write_hit(hits, i, t, tri4, ntri);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Persistent mapping over trees with at least two leaves.  queue[0] is the
// next unclaimed ray index, queue[1] counts finished waves; both are zero at
// launch and are zero again when the last wave exits.
template <bool COUNT>
__global__ __launch_bounds__(kRayBlock) void cast_rays_persistent_kernel(
    const pq_bvh_node_wide *__restrict__ nodes, const TriRecord *__restrict__ tris, uint32_t ntri,
    const pq_ray *__restrict__ rays, uint64_t n, float max_toi, pq_ray_hit *__restrict__ hits,
    uint32_t *__restrict__ counts, unsigned long long *__restrict__ queue, uint32_t chunk) {
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stk = lds_stack + threadIdx.x;
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);
    const uint32_t lane = __lane_id();

    uint64_t q_next = 0, q_end = 0;  // this wave's claimed range [q_next, q_end)
    bool drained = false;            // the shared queue is empty (wave-uniform)
    bool active = false;
    uint64_t ray_id = 0;
    LaneRay r{};
    Traversal t;
    start_traversal(t, max_toi);
    uint32_t n_wide = 0, n_tri = 0;

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (idle != 0 && !drained) {
            if (q_next >= q_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&queue[0], (unsigned long long)chunk);
                base = __shfl(base, 0);
                q_next = base;
                q_end = base + chunk < n ? base + chunk : n;
                drained = base >= n;
            }
            if (!drained) {
                const uint64_t avail = q_end - q_next;
                const uint32_t rank = lanes_below(idle);
                if (!active && rank < avail) {
                    ray_id = q_next + rank;
                    r = load_ray(rays, ray_id);
                    start_traversal(t, max_toi);
                    n_wide = 0;
                    n_tri = 0;
                    active = true;
                }
                const uint64_t took = (uint64_t)__popcll(idle);
                q_next += took < avail ? took : avail;
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active && find_best_step<COUNT>(node4, tri4, r, t, stk, n_wide, n_tri)) {
            if (COUNT) {
                counts[2 * ray_id] = n_wide;
                counts[2 * ray_id + 1] = n_tri;
            }
            write_hit(hits, ray_id, t, tri4, ntri);
            active = false;
        }
    }
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (kRayBlock / 64);
        const unsigned long long done = atomicAdd(&queue[1], 1ull);
        if (done == waves - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Batched casts over many meshes in one launch.  Rays are concatenated mesh by
// mesh; ray i belongs to the mesh whose [offset, next offset) range holds it
// (binary search over the offsets, which stay in cache).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mesh_of(const unsigned long long *offsets, uint32_t nmesh, uint64_t i) {
    uint32_t lo = 0, hi = nmesh;  // offsets[lo] <= i < offsets[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <bool COUNT>
__device__ __forceinline__ void cast_one(const BatchMesh &bm, const LaneRay &r, float max_toi, uint32_t *stk,
                                         Traversal &t, uint32_t &n_wide, uint32_t &n_tri) {
    const float4 *node4 = reinterpret_cast<const float4 *>(bm.nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(bm.tris);
    start_traversal(t, max_toi);
    n_wide = n_tri = 0;
    if (bm.single_leaf) {
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
            visit_leaf<true>(tri4, __float_as_uint(q0.w), r, t, n_tri);
        return;
    }
    while (!find_best_step<COUNT>(node4, tri4, r, t, stk, n_wide, n_tri)) {
    }
}

template <bool COUNT>
__global__ __launch_bounds__(kRayBlock) void cast_rays_batch_kernel(const BatchMesh *__restrict__ meshes,
                                                                    const unsigned long long *__restrict__ offsets,
                                                                    uint32_t nmesh, const pq_ray *__restrict__ rays,
                                                                    uint64_t n, float max_toi,
                                                                    pq_ray_hit *__restrict__ hits,
                                                                    uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t i = (uint64_t)blockIdx.x * kRayBlock + threadIdx.x;
    if (i >= n) return;
    const BatchMesh bm = meshes[mesh_of(offsets, nmesh, i)];
    const LaneRay r = load_ray(rays, i);
    Traversal t;
    uint32_t n_wide, n_tri;
    cast_one<COUNT>(bm, r, max_toi, lds_stack + threadIdx.x, t, n_wide, n_tri);
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }
    write_hit(hits, i, t, reinterpret_cast<const float4 *>(bm.tris), bm.ntri);
}

// Persistent variant: one device queue over the concatenated rays; finished
// lanes take the next rays every iteration, whatever mesh they belong to.
template <bool COUNT>
__global__ __launch_bounds__(kRayBlock) void cast_rays_batch_persistent_kernel(
    const BatchMesh *__restrict__ meshes, const unsigned long long *__restrict__ offsets, uint32_t nmesh,
    const pq_ray *__restrict__ rays, uint64_t n, float max_toi, pq_ray_hit *__restrict__ hits,
    uint32_t *__restrict__ counts, unsigned long long *__restrict__ queue, uint32_t chunk) {
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stk = lds_stack + threadIdx.x;
    const uint32_t lane = __lane_id();
    uint64_t q_next = 0, q_end = 0;
    bool drained = false, active = false;
    uint64_t ray_id = 0;
    LaneRay r{};
    BatchMesh bm{};
    Traversal t;
    start_traversal(t, max_toi);
    uint32_t n_wide = 0, n_tri = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (idle != 0 && !drained) {
            if (q_next >= q_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&queue[0], (unsigned long long)chunk);
                base = __shfl(base, 0);
                q_next = base;
                q_end = base + chunk < n ? base + chunk : n;
                drained = base >= n;
            }
            if (!drained) {
                const uint64_t avail = q_end - q_next;
                const uint32_t rank = lanes_below(idle);
                if (!active && rank < avail) {
                    ray_id = q_next + rank;
                    bm = meshes[mesh_of(offsets, nmesh, ray_id)];
                    r = load_ray(rays, ray_id);
                    start_traversal(t, max_toi);
                    n_wide = n_tri = 0;
                    active = true;
                    if (bm.single_leaf) {  // one-triangle mesh: a single test
                        const float4 *node4 = reinterpret_cast<const float4 *>(bm.nodes);
                        const float4 q0 = node4[0], q1 = node4[1];
                        n_wide = 1;
                        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
                            visit_leaf<true>(reinterpret_cast<const float4 *>(bm.tris), __float_as_uint(q0.w), r, t,
                                             n_tri);
                    }
                }
                const uint64_t took = (uint64_t)__popcll(idle);
                q_next += took < avail ? took : avail;
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active) {
            const bool done = bm.single_leaf ||
                              find_best_step<COUNT>(reinterpret_cast<const float4 *>(bm.nodes),
                                                    reinterpret_cast<const float4 *>(bm.tris), r, t, stk, n_wide, n_tri);
            if (done) {
                if (COUNT) {
                    counts[2 * ray_id] = n_wide;
                    counts[2 * ray_id + 1] = n_tri;
                }
                write_hit(hits, ray_id, t, reinterpret_cast<const float4 *>(bm.tris), bm.ntri);
                active = false;
            }
        }
    }
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (kRayBlock / 64);
        const unsigned long long done = atomicAdd(&queue[1], 1ull);
        if (done == waves - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

}  // namespace

uint32_t ray_stack_entries(uint32_t depth) { return depth > 1 ? depth - 1 : 1; }

hipError_t launch_cast_rays_culling(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi,
                                    uint32_t culling, pq_ray_hit *hits, uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t lds = (size_t)ray_stack_entries(m.depth) * kRayBlock * sizeof(uint32_t);
    const dim3 grid((uint32_t)((n + kRayBlock - 1) / kRayBlock)), block(kRayBlock);
    if (culling == 1 && counts)
        hipLaunchKernelGGL((cast_rays_kernel<true, 1>), grid, block, lds, stream, m.nodes, m.tris, m.num_triangles,
                           rays, n, max_toi, hits, counts);
    else if (culling == 1)
        hipLaunchKernelGGL((cast_rays_kernel<false, 1>), grid, block, lds, stream, m.nodes, m.tris, m.num_triangles,
                           rays, n, max_toi, hits, counts);
    else if (counts)
        hipLaunchKernelGGL((cast_rays_kernel<true, 2>), grid, block, lds, stream, m.nodes, m.tris, m.num_triangles,
                           rays, n, max_toi, hits, counts);
    else
        hipLaunchKernelGGL((cast_rays_kernel<false, 2>), grid, block, lds, stream, m.nodes, m.tris, m.num_triangles,
                           rays, n, max_toi, hits, counts);
    return hipGetLastError();
}

hipError_t launch_cast_rays(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi, pq_ray_hit *hits,
                            uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t lds = (size_t)ray_stack_entries(m.depth) * kRayBlock * sizeof(uint32_t);
    const bool single_leaf = m.num_nodes == 1 && m.num_triangles == 1;
    if (m.tuning.mapping == PQ_RAY_MAPPING_PERSISTENT && !single_leaf) {
        unsigned long long *queue = nullptr;
        hipError_t e = stream_queue(stream, &queue);
        if (e != hipSuccess) return e;
        int dev = 0, cus = 0, per_cu = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        if (counts)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cast_rays_persistent_kernel<true>, kRayBlock, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cast_rays_persistent_kernel<false>, kRayBlock, lds);
        if (e != hipSuccess) return e;
        if (m.tuning.blocks_per_cu > 0 && (int)m.tuning.blocks_per_cu < per_cu) per_cu = (int)m.tuning.blocks_per_cu;
        if (per_cu < 1) per_cu = 1;
        const uint32_t chunk = m.tuning.chunk ? m.tuning.chunk : 128;
        uint64_t want = (n + chunk - 1) / chunk;  // no more waves than chunks
        uint64_t blocks = (uint64_t)cus * (uint64_t)per_cu;
        const uint64_t waves_per_block = kRayBlock / 64;
        if (blocks * waves_per_block > want) blocks = (want + waves_per_block - 1) / waves_per_block;
        if (blocks < 1) blocks = 1;
        if (counts)
            hipLaunchKernelGGL(cast_rays_persistent_kernel<true>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream,
                               m.nodes, m.tris, m.num_triangles, rays, n, max_toi, hits, counts, queue, chunk);
        else
            hipLaunchKernelGGL(cast_rays_persistent_kernel<false>, dim3((uint32_t)blocks), dim3(kRayBlock), lds,
                               stream, m.nodes, m.tris, m.num_triangles, rays, n, max_toi, hits, counts, queue, chunk);
        return hipGetLastError();
    }
    const uint64_t blocks = (n + kRayBlock - 1) / kRayBlock;
    if (counts)
        hipLaunchKernelGGL(cast_rays_kernel<true>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream, m.nodes,
                           m.tris, m.num_triangles, rays, n, max_toi, hits, counts);
    else
        hipLaunchKernelGGL(cast_rays_kernel<false>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream, m.nodes,
                           m.tris, m.num_triangles, rays, n, max_toi, hits, counts);
    return hipGetLastError();
}

hipError_t launch_cast_rays_batch(const DeviceBatch &b, const pq_ray *rays, float max_toi, pq_ray_hit *hits,
                                  uint32_t *counts, hipStream_t stream) {
    const uint64_t n = b.num_rays;
    if (n == 0) return hipSuccess;
    const size_t lds = (size_t)ray_stack_entries(b.max_depth) * kRayBlock * sizeof(uint32_t);
    if (b.tuning.mapping == PQ_RAY_MAPPING_PERSISTENT) {
        hipError_t e;
        int dev = 0, cus = 0, per_cu = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        if (counts)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cast_rays_batch_persistent_kernel<true>,
                                                             kRayBlock, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cast_rays_batch_persistent_kernel<false>,
                                                             kRayBlock, lds);
        if (e != hipSuccess) return e;
        if (b.tuning.blocks_per_cu > 0 && (int)b.tuning.blocks_per_cu < per_cu) per_cu = (int)b.tuning.blocks_per_cu;
        if (per_cu < 1) per_cu = 1;
        const uint32_t chunk = b.tuning.chunk ? b.tuning.chunk : 128;
        const uint64_t want = (n + chunk - 1) / chunk;
        uint64_t blocks = (uint64_t)cus * (uint64_t)per_cu;
        if (blocks * (kRayBlock / 64) > want) blocks = (want + (kRayBlock / 64) - 1) / (kRayBlock / 64);
        if (blocks < 1) blocks = 1;
        if (counts)
            hipLaunchKernelGGL(cast_rays_batch_persistent_kernel<true>, dim3((uint32_t)blocks), dim3(kRayBlock), lds,
                               stream, b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts, b.queue,
                               chunk);
        else
            hipLaunchKernelGGL(cast_rays_batch_persistent_kernel<false>, dim3((uint32_t)blocks), dim3(kRayBlock), lds,
                               stream, b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts, b.queue,
                               chunk);
        return hipGetLastError();
    }
    const uint64_t blocks = (n + kRayBlock - 1) / kRayBlock;
    if (counts)
        hipLaunchKernelGGL(cast_rays_batch_kernel<true>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream,
                           b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts);
    else
        hipLaunchKernelGGL(cast_rays_batch_kernel<false>, dim3((uint32_t)blocks), dim3(kRayBlock), lds, stream,
                           b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts);
    return hipGetLastError();
}

}  // namespace pq
