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
// loads.  The pending-node stack (at most depth-1 entries) keeps its top
// entries in registers and the rest in LDS, strided by the block size so each
// lane's slots are bank-conflict free.
//
// Four kernel mappings (pq_ray_tuning, chosen per mesh or per batch by the
// tuner; every one gives the same hits and the same node and triangle counts):
//   * static: one launch lane per ray; a wave lasts as long as its slowest ray;
//   * binned: static lanes over the rays counting-sorted by the cell where they
//     enter the mesh's bounding region (see the sort below), so the rays of a
//     wave walk similar paths and share cached nodes; in a batch each mesh's
//     rays are sorted on their own and stay in their mesh's range;
//   * persistent: a grid sized to the resident waves; each wave pulls chunks
//     of ray indices from a queue in device memory, and every iteration the
//     lanes whose ray has finished (found with a wave ballot) take the next
//     rays of the chunk, ranked by a masked popcount, so the wave stays full
//     until the queue drains.  The queue counter resets itself: the last
//     wave to finish zeroes it for the next launch on the stream;
//   * wave: one wave per ray, 64 wide nodes fetched at once (see below).
// The lane mappings are instantiated for 64-, 128- and 256-lane workgroups
// and 0, 1, 2 or 4 stack entries in registers.
//
// Two exact shortcuts in the static and binned mappings: a wave whose rays
// have no zero direction component (and finite origins and inverses, on a mesh
// with finite bounds) runs a slab test without the degenerate-axis terms, and
// the traversal keeps only the best time of impact, leaf and side, computing
// the winner's normal once at the end.
//
// Floating point: compiled with -ffp-contract=off; 1/x and sqrt are the
// correctly rounded IEEE operations, min/max are IEEE maxNum/minNum (the
// semantics of Rust's f32::min/max), so every value is the one parry computes.
#include <cfloat>
#include <type_traits>

#include "pq_internal.h"

namespace pq {
namespace {

struct LaneRay {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;  // 1/dir, the `denom` of ray_aabb.rs:22 hoisted out of the node loop
};

// Aabb::cast_local_ray with solid = true, mapped to BvhNode::cast_ray's
// unwrap_or(Real::MAX).  The reference exits as soon as tmin > tmax; because
// tmin only grows and tmax only shrinks, testing once after the three axes
// gives the same answer, so the three axes are evaluated branch-free.
//
// REGULAR: the caller guarantees dx, dy and dz are nonzero, the origin and
// 1/dir finite, and the node bounds finite.  Then every axis takes the slab
// branch and the degenerate-axis terms are identities, left out; and no
// product (finite bound - finite origin) * finite inverse is NaN, so ordering
// the two slab ends by select equals taking min and max (they can differ only
// in the sign of a zero end, which no comparison of the cost distinguishes).
template <bool REGULAR = false>
__device__ __forceinline__ float slab_cost(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                           const LaneRay &r, float max_toi) {
    float tmin = 0.0f, tmax = max_toi;
    // Bitwise (not short-circuit) boolean algebra keeps the test free of
    // exec-mask branches: every operand is a plain compare.
    unsigned miss = 0u;
#define PQ_SLAB_AXIS(MN, MX, O, D, INV)                               \
    {                                                                 \
        const float n = (MN - O) * INV;                                \
        const float f = (MX - O) * INV;                                \
        float lo, hi;                                                  \
        if constexpr (REGULAR) {                                       \
            lo = fminf(n, f);                                          \
            hi = fmaxf(n, f);                                          \
        } else {                                                       \
            const bool s = n > f;                                      \
            lo = s ? f : n;                                            \
            hi = s ? n : f;                                            \
            const unsigned z = (unsigned)(D == 0.0f);                  \
            miss |= z & ((unsigned)(O < MN) | (unsigned)(O > MX));     \
            lo = z ? -INFINITY : lo;                                   \
            hi = z ? INFINITY : hi;                                    \
        }                                                              \
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
// the `toi <= max_toi` acceptance of Triangle::cast_local_ray_and_get_normal,
// without the normal: only the time of impact and the side decide find_best's
// path, so the traversal defers the normal to the winning triangle
// (triangle_normal below, the same operations on the same values).
// CULL 1 / 2 first drops the triangles RayCullingMode::IgnoreBackfaces /
// IgnoreFrontfaces skips (ray_trimesh.rs:58-65): its test is the sign of the
// scaled normal (ab x ac) dotted with the ray direction, which is `d` below.
template <int CULL = 0>
__device__ __forceinline__ bool ray_triangle_toi(const float4 q0, const float4 q1, const float4 q2, const LaneRay &r,
                                                 float max_toi, float &toi_out, uint32_t &back_out) {
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
    if (!(toi <= max_toi)) return false;
    toi_out = toi;
    back_out = back;
    return true;
}

// The unit normal ray_triangle.rs:70-152 reports with a hit of this ray on
// this triangle: (ab x ac) normalised, negated when the origin lies below the
// triangle's plane.  Recomputes n and t exactly as ray_triangle_toi does.
__device__ __forceinline__ void triangle_normal(const float4 q0, const float4 q1, const float4 q2, const LaneRay &r,
                                                float &ux, float &uy, float &uz) {
    const float ax = q0.x, ay = q0.y, az = q0.z;
    const float bx = q0.w, by = q1.x, bz = q1.y;
    const float cx = q1.z, cy = q1.w, cz = q2.x;
    const float abx = bx - ax, aby = by - ay, abz = bz - az;
    const float acx = cx - ax, acy = cy - ay, acz = cz - az;
    const float nx = aby * acz - abz * acy;
    const float ny = abz * acx - abx * acz;
    const float nz = abx * acy - aby * acx;
    const float apx = r.ox - ax, apy = r.oy - ay, apz = r.oz - az;
    const float t = (apx * nx + apy * ny) + apz * nz;
    const float len = sqrtf((nx * nx + ny * ny) + nz * nz);
    ux = nx / len;
    uy = ny / len;
    uz = nz / len;
    if (t < 0.0f) {
        ux = -ux;
        uy = -uy;
        uz = -uz;
    }
}

// Time of impact, side and normal together (a single test, no traversal).
template <int CULL = 0>
__device__ __forceinline__ bool ray_triangle(const float4 q0, const float4 q1, const float4 q2, const LaneRay &r,
                                             float max_toi, TriHit &h) {
    if (!ray_triangle_toi<CULL>(q0, q1, q2, r, max_toi, h.toi, h.back)) return false;
    triangle_normal(q0, q1, q2, r, h.nx, h.ny, h.nz);
    return true;
}

__device__ __forceinline__ bool node_is_leaf(uint32_t data) { return (data & 0x3fffffffu) == 1u; }

// The pending-node stack of find_best (bvh_traverse.rs:361-416), one per lane:
// the top REG entries in registers, the rest in this lane's LDS column (stride
// BLOCK words, so the lanes of a wave hit distinct banks).  The register slots
// shift by value with constant indices, so they never go to scratch.  A pop
// refills the deepest register slot from LDS, a load nothing waits on until
// REG more pops.  It never holds more than depth-1 entries (each pending entry
// sits one level deeper than the one below it); the host sizes the LDS part to
// that less REG.
template <int BLOCK, int REG>
struct LaneStack {
    uint32_t *col;
    uint32_t reg[REG > 0 ? REG : 1];
    uint32_t sp;
    __device__ __forceinline__ void reset() { sp = 0; }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ void push(uint32_t x) {
        if constexpr (REG == 0) {
            col[sp * BLOCK] = x;
        } else {
            if (sp >= (uint32_t)REG) col[(sp - REG) * BLOCK] = reg[REG - 1];
#pragma unroll
            for (int k = REG - 1; k > 0; --k) reg[k] = reg[k - 1];
            reg[0] = x;
        }
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if constexpr (REG == 0) {
            return col[sp * BLOCK];
        } else {
            const uint32_t x = reg[0];
#pragma unroll
            for (int k = 0; k < REG - 1; ++k) reg[k] = reg[k + 1];
            if (sp >= (uint32_t)REG) reg[REG - 1] = col[(sp - REG) * BLOCK];
            return x;
        }
    }
};

// A lane's find_best state: the best time of impact so far (max_toi before
// any hit), the leaf holding it and its side; the normal is computed once,
// from the winning triangle, when the hit is written.
struct Traversal {
    float best_cost;
    uint32_t best_leaf;
    uint32_t best_back;
    uint32_t curr;
    bool have;
    float nx, ny, nz;  // the best hit's normal, kept only when it is not deferred
};

// A ray record as loaded (origin, direction), before the inverses.
struct RawRay {
    float2 r01, r23, r45;
};

__device__ __forceinline__ RawRay load_raw_ray(const pq_ray *rays, uint64_t i) {
    const float *rp = reinterpret_cast<const float *>(rays + i);
    return RawRay{*reinterpret_cast<const float2 *>(rp), *reinterpret_cast<const float2 *>(rp + 2),
                  *reinterpret_cast<const float2 *>(rp + 4)};
}

__device__ __forceinline__ LaneRay lane_ray(const RawRay &w) {
    LaneRay r;
    r.ox = w.r01.x;
    r.oy = w.r01.y;
    r.oz = w.r23.x;
    r.dx = w.r23.y;
    r.dy = w.r45.x;
    r.dz = w.r45.y;
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;
    return r;
}

__device__ __forceinline__ LaneRay load_ray(const pq_ray *rays, uint64_t i) { return lane_ray(load_raw_ray(rays, i)); }

// No zero direction component, a finite origin and finite inverses:
// slab_cost<true> applies to this ray on a mesh with finite bounds.
__device__ __forceinline__ bool ray_is_regular(const LaneRay &r) {
    return r.dx != 0.0f && r.dy != 0.0f && r.dz != 0.0f && isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) &&
           isfinite(r.ix) && isfinite(r.iy) && isfinite(r.iz);
}

__device__ __forceinline__ void start_traversal(Traversal &t, float max_toi) {
    t.best_cost = max_toi;
    t.best_leaf = 0;
    t.best_back = 0;
    t.curr = 0;
    t.have = false;
    t.nx = t.ny = t.nz = 0.0f;
}

// Tests one leaf (a triangle record) against the current best.  DEFER leaves
// the normal to write_hit; otherwise it is computed with each new best.
template <bool COUNT, int CULL = 0, bool DEFER = true>
__device__ __forceinline__ void visit_leaf(const float4 *tri4, uint32_t leaf, const LaneRay &r, Traversal &t,
                                           uint32_t &n_tri) {
    float toi;
    uint32_t back;
    if (COUNT) ++n_tri;
    const float4 q0 = tri4[3 * leaf], q1 = tri4[3 * leaf + 1], q2 = tri4[3 * leaf + 2];
    if (ray_triangle_toi<CULL>(q0, q1, q2, r, t.best_cost, toi, back) && toi < t.best_cost) {
        t.best_cost = toi;
        t.best_back = back;
        t.best_leaf = leaf;
        t.have = true;
        if constexpr (!DEFER) triangle_normal(q0, q1, q2, r, t.nx, t.ny, t.nz);
    }
}

// One iteration of find_best's loop (bvh_traverse.rs:361-416) on wide node
// t.curr.  Returns true when the traversal has finished.
template <bool COUNT, int CULL, bool REGULAR, bool DEFER = true, class Stack>
__device__ __forceinline__ bool find_best_step(const float4 *node4, const float4 *tri4, const LaneRay &r,
                                               Traversal &t, Stack &stk, uint32_t &n_wide, uint32_t &n_tri) {
    if (COUNT) ++n_wide;
    const float4 q0 = node4[4 * t.curr + 0];  // left mins, left children
    const float4 q1 = node4[4 * t.curr + 1];  // left maxs, left data
    const float4 q2 = node4[4 * t.curr + 2];  // right mins, right children
    const float4 q3 = node4[4 * t.curr + 3];  // right maxs, right data
    float ls = slab_cost<REGULAR>(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, t.best_cost);
    float rs = slab_cost<REGULAR>(q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, r, t.best_cost);
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
            visit_leaf<COUNT, CULL, DEFER>(tri4, lc, r, t, n_tri);
        } else {
            t.curr = lc;
            found_next = true;
        }
    }
    if (rs < t.best_cost && rs != FLT_MAX) {
        if (node_is_leaf(rd)) {
            visit_leaf<COUNT, CULL, DEFER>(tri4, rc, r, t, n_tri);
        } else if (found_next) {
            stk.push(rc);
        } else {
            t.curr = rc;
            found_next = true;
        }
    }
    if (!found_next) {
        if (stk.empty()) return true;
        t.curr = stk.pop();
    }
    return false;
}

__device__ __forceinline__ pq_ray_hit make_hit(bool have, const TriHit &h, uint32_t tri_id, uint32_t ntri) {
    pq_ray_hit out;
    if (have) {
        out.time_of_impact = h.toi;
        out.normal[0] = h.nx;
        out.normal[1] = h.ny;
        out.normal[2] = h.nz;
        out.feature = h.back ? tri_id + ntri : tri_id;  // ray_trimesh.rs:25-33
    } else {
        out.time_of_impact = 0.0f;
        out.normal[0] = out.normal[1] = out.normal[2] = 0.0f;
        out.feature = PQ_NO_HIT;
    }
    return out;
}

template <bool DEFER = true>
__device__ __forceinline__ void write_hit(pq_ray_hit *hits, uint64_t i, const Traversal &t, const LaneRay &r,
                                          const float4 *tri4, uint32_t ntri) {
    TriHit h{};
    uint32_t tri_id = 0;
    if (t.have) {
        const float4 q2 = tri4[3 * t.best_leaf + 2];
        if constexpr (DEFER) {
            triangle_normal(tri4[3 * t.best_leaf], tri4[3 * t.best_leaf + 1], q2, r, h.nx, h.ny, h.nz);
        } else {
            h.nx = t.nx;
            h.ny = t.ny;
            h.nz = t.nz;
        }
        h.toi = t.best_cost;
        h.back = t.best_back;
        tri_id = __float_as_uint(q2.y);
    }
    hits[i] = make_hit(t.have, h, tri_id, ntri);
}

// Static mapping; also serves trees whose root holds a single leaf
// (bvh_traverse.rs:348-359).  COUNT = true additionally writes, per ray, the
// number of wide-node records fetched and of triangles tested (the traffic
// model of the roofline figure).  CULL selects the culling variant
// (TriMesh::cast_local_ray_with_culling, ray_trimesh.rs:152-178).
template <bool COUNT, int CULL, int BLOCK, int REG>
__global__ __launch_bounds__(BLOCK) void cast_rays_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                          const TriRecord *__restrict__ tris, uint32_t ntri,
                                                          const pq_ray *__restrict__ rays, uint64_t n, float max_toi,
                                                          pq_ray_hit *__restrict__ hits,
                                                          uint32_t *__restrict__ counts,
                                                          const uint32_t *__restrict__ order, uint32_t finite) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const uint64_t i = order ? (uint64_t)order[j] : j;  // the binned mapping's j-th ray in sorted order
    LaneStack<BLOCK, REG> stk;
    stk.col = lds_stack + threadIdx.x;
    stk.reset();
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
    } else if (finite && __all(ray_is_regular(r))) {  // the usual case: the whole wave takes the short test
        while (!find_best_step<COUNT, CULL, true>(node4, tri4, r, t, stk, n_wide, n_tri)) {
        }
    } else {
        while (!find_best_step<COUNT, CULL, false>(node4, tri4, r, t, stk, n_wide, n_tri)) {
        }
    }
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }
    write_hit(hits, i, t, r, tri4, ntri);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The last wave of a persistent grid to finish zeroes the queue for the next
// launch on the stream.  queue[0] is the next unclaimed ray index, queue[1]
// counts finished waves.
__device__ __forceinline__ void queue_wave_done(unsigned long long *queue, uint32_t waves) {
    if (__lane_id() == 0) {
        __threadfence();
        const unsigned long long done = atomicAdd(&queue[1], 1ull);
        if (done == waves - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

// Persistent lane mapping over trees with at least two leaves.  It runs the
// general slab test and computes each new best's normal at once: a lane that
// finishes writes its hit inside the loop, in a block of its own, so a
// deferred normal would cost that block every iteration, and a second,
// regular copy of the step raises register use; both were slower here.
template <bool COUNT, int BLOCK, int REG>
__global__ __launch_bounds__(BLOCK) void cast_rays_persistent_kernel(
    const pq_bvh_node_wide *__restrict__ nodes, const TriRecord *__restrict__ tris, uint32_t ntri,
    const pq_ray *__restrict__ rays, uint64_t n, float max_toi, pq_ray_hit *__restrict__ hits,
    uint32_t *__restrict__ counts, unsigned long long *__restrict__ queue, uint32_t chunk) {
    extern __shared__ uint32_t lds_stack[];
    LaneStack<BLOCK, REG> stk;
    stk.col = lds_stack + threadIdx.x;
    stk.reset();
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);

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
                if (__lane_id() == 0) base = atomicAdd(&queue[0], (unsigned long long)chunk);
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
                    stk.reset();
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
        if (active && find_best_step<COUNT, 0, false, false>(node4, tri4, r, t, stk, n_wide, n_tri)) {
            if (COUNT) {
                counts[2 * ray_id] = n_wide;
                counts[2 * ray_id + 1] = n_tri;
            }
            write_hit<false>(hits, ray_id, t, r, tri4, ntri);
            active = false;
        }
    }
    queue_wave_done(queue, gridDim.x * (BLOCK / 64));
}

// ---------------------------------------------------------------------------
// Wave per ray.  find_best's decisions only compare a child's slab cost and a
// leaf's time of impact with the current best, so both can be computed ahead
// against max_toi instead of the best of the moment:
//   * slab test: with tmax = best the cost is tmin if tmin <= min(best, exit)
//     and f32::MAX otherwise; with tmax = max_toi (>= best) it is tmin in more
//     cases, but only where tmin >= best, i.e. for children that are never
//     entered either way, and ordering the two children by the max_toi costs
//     puts an enterable child first exactly when the best-bounded costs do;
//   * triangle: `toi <= best && toi < best` is `toi <= max_toi && toi < best`.
// So a wave fetches a window of 64 consecutive wide nodes (in depth-first
// order, the first 64 nodes of the subtree under the window's first node),
// lane k computing both children of node base + k, including the triangle of
// each leaf child it could enter, and then runs find_best's loop with uniform
// (scalar) state, reading node k's results from lane k.  A node outside the
// window starts a new window there.  The pending-node stack is one VGPR: entry
// k in lane k (depth-1 <= 64 entries; the host keeps deeper trees on the lane
// mappings).  The winner's normal is recomputed from its triangle at the end.
// ---------------------------------------------------------------------------
struct WaveWindow {
    float ls, rs, lt, rt;   // child slab costs and leaf times of impact (+inf: no hit)
    uint32_t lc, rc, ld, rd;
};

template <bool REGULAR>
__device__ __forceinline__ WaveWindow fetch_window(const float4 *node4, uint32_t num_nodes, const float4 *tri4,
                                                   uint32_t base, const LaneRay &r, float max_toi) {
    WaveWindow w{FLT_MAX, FLT_MAX, INFINITY, INFINITY, 0u, 0u, 0u, 0u};
    const uint32_t k = base + __lane_id();
    if (k < num_nodes) {
        const float4 q0 = node4[4 * k + 0], q1 = node4[4 * k + 1];
        const float4 q2 = node4[4 * k + 2], q3 = node4[4 * k + 3];
        w.ls = slab_cost<REGULAR>(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi);
        w.rs = slab_cost<REGULAR>(q2.x, q2.y, q2.z, q3.x, q3.y, q3.z, r, max_toi);
        w.lc = __float_as_uint(q0.w);
        w.ld = __float_as_uint(q1.w);
        w.rc = __float_as_uint(q2.w);
        w.rd = __float_as_uint(q3.w);
        float toi;
        uint32_t back;
        if (node_is_leaf(w.ld) && w.ls < max_toi && w.ls != FLT_MAX &&
            ray_triangle_toi<0>(tri4[3 * w.lc], tri4[3 * w.lc + 1], tri4[3 * w.lc + 2], r, max_toi, toi, back))
            w.lt = toi;
        if (node_is_leaf(w.rd) && w.rs < max_toi && w.rs != FLT_MAX &&
            ray_triangle_toi<0>(tri4[3 * w.rc], tri4[3 * w.rc + 1], tri4[3 * w.rc + 2], r, max_toi, toi, back))
            w.rt = toi;
    }
    return w;
}

__device__ __forceinline__ float lane_f(float v, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ uint32_t lane_u(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// find_best for one ray by the whole wave; every lane returns the hit.
template <bool COUNT>
__device__ __forceinline__ pq_ray_hit wave_cast_one(const float4 *node4, uint32_t num_nodes, const float4 *tri4,
                                                    uint32_t finite,
                                                    uint32_t ntri, bool single_leaf, const LaneRay &r,
                                                    float max_toi, uint32_t &n_wide, uint32_t &n_tri) {
    float best = max_toi;
    uint32_t best_leaf = 0;
    bool have = false;
    n_wide = n_tri = 0;
    if (single_leaf) {
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        n_tri = 0;
        float toi;
        uint32_t back;
        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi) {
            n_tri = 1;
            const uint32_t leaf = __float_as_uint(q0.w);
            if (ray_triangle_toi<0>(tri4[3 * leaf], tri4[3 * leaf + 1], tri4[3 * leaf + 2], r, max_toi, toi, back) &&
                toi < best) {
                best_leaf = leaf;
                have = true;
            }
        }
    } else {
        uint32_t stk = 0, sp = 0, curr = 0, base = 0;
        const bool regular = finite && ray_is_regular(r);  // the same ray in every lane: a uniform branch
        auto fetch = [&](uint32_t at) {
            return regular ? fetch_window<true>(node4, num_nodes, tri4, at, r, max_toi)
                           : fetch_window<false>(node4, num_nodes, tri4, at, r, max_toi);
        };
        WaveWindow w = fetch(0u);
        for (;;) {
            if (curr - base >= 64u) {
                base = curr;
                w = fetch(base);
            }
            const uint32_t j = curr - base;
            if (COUNT) ++n_wide;
            float ls = lane_f(w.ls, j), rs = lane_f(w.rs, j);
            float lt = lane_f(w.lt, j), rt = lane_f(w.rt, j);
            uint32_t lc = lane_u(w.lc, j), rc = lane_u(w.rc, j);
            uint32_t ld = lane_u(w.ld, j), rd = lane_u(w.rd, j);
            if (ls > rs) {
                float tf = ls;
                ls = rs;
                rs = tf;
                tf = lt;
                lt = rt;
                rt = tf;
                uint32_t tu = lc;
                lc = rc;
                rc = tu;
                tu = ld;
                ld = rd;
                rd = tu;
            }
            bool found_next = false;
            if (ls < best && ls != FLT_MAX) {
                if (node_is_leaf(ld)) {
                    if (COUNT) ++n_tri;
                    if (lt < best) {
                        best = lt;
                        best_leaf = lc;
                        have = true;
                    }
                } else {
                    curr = lc;
                    found_next = true;
                }
            }
            if (rs < best && rs != FLT_MAX) {
                if (node_is_leaf(rd)) {
                    if (COUNT) ++n_tri;
                    if (rt < best) {
                        best = rt;
                        best_leaf = rc;
                        have = true;
                    }
                } else if (found_next) {
                    stk = __lane_id() == sp ? rc : stk;
                    ++sp;
                } else {
                    curr = rc;
                    found_next = true;
                }
            }
            if (!found_next) {
                if (sp == 0) break;
                --sp;
                curr = lane_u(stk, sp);
            }
        }
    }
    TriHit h{};
    uint32_t tri_id = 0;
    if (have) {
        (void)ray_triangle<0>(tri4[3 * best_leaf], tri4[3 * best_leaf + 1], tri4[3 * best_leaf + 2], r, max_toi, h);
        tri_id = __float_as_uint(tri4[3 * best_leaf + 2].y);
    }
    return make_hit(have, h, tri_id, ntri);
}

// Claims the next `chunk` rays of the queue for the whole wave: [*begin, *end).
__device__ __forceinline__ bool wave_claim(unsigned long long *queue, uint32_t chunk, uint64_t n, uint64_t &begin,
                                           uint64_t &end) {
    unsigned long long b = 0;
    if (__lane_id() == 0) b = atomicAdd(&queue[0], (unsigned long long)chunk);
    b = __builtin_amdgcn_readfirstlane((uint32_t)b) | ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(b >> 32)) << 32);
    begin = b;
    end = b + chunk < n ? b + chunk : n;
    return b < n;
}

template <bool COUNT>
__global__ __launch_bounds__(64) void cast_rays_wave_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                            uint32_t num_nodes, const TriRecord *__restrict__ tris,
                                                            uint32_t ntri, const pq_ray *__restrict__ rays, uint64_t n,
                                                            float max_toi, pq_ray_hit *__restrict__ hits,
                                                            uint32_t *__restrict__ counts,
                                                            unsigned long long *__restrict__ queue, uint32_t chunk,
                                                            uint32_t finite) {
    const float4 *node4 = reinterpret_cast<const float4 *>(nodes);
    const float4 *tri4 = reinterpret_cast<const float4 *>(tris);
    const bool single_leaf = (__float_as_uint(node4[3].w) & 0x3fffffffu) == 0u;
    uint64_t begin, end;
    while (wave_claim(queue, chunk, n, begin, end)) {
        for (uint64_t i = begin; i < end; ++i) {
            const LaneRay r = load_ray(rays, i);
            uint32_t n_wide, n_tri;
            const pq_ray_hit h = wave_cast_one<COUNT>(node4, num_nodes, tri4, finite, ntri, single_leaf, r, max_toi,
                                                      n_wide,
                                                      n_tri);
            if (__lane_id() == 0) {
                hits[i] = h;
                if (COUNT) {
                    counts[2 * i] = n_wide;
                    counts[2 * i + 1] = n_tri;
                }
            }
        }
    }
    queue_wave_done(queue, gridDim.x);
}

// ---------------------------------------------------------------------------
// Batched casts over many meshes in one launch.  Rays are concatenated mesh by
// mesh; ray i belongs to the mesh whose [offset, next offset) range holds it
// (binary search over the offsets, which stay in cache).
// ---------------------------------------------------------------------------
// A mesh's arrays are device global memory, but the compiler takes pointers read
// from the BatchMesh table as generic and fetches through them with flat
// loads, which also count against the LDS wait counter.  Rebuilding the
// pointer from its address in the global address space makes them global
// loads (a plain cast of the generic pointer is folded away).
__device__ __forceinline__ const float4 *global_float4(const void *p) {
    typedef const __attribute__((address_space(1))) float4 *global_ptr;
    return (const float4 *)(global_ptr)(uintptr_t)p;
}

__device__ __forceinline__ uint32_t mesh_of(const unsigned long long *offsets, uint32_t nmesh, uint64_t i) {
    uint32_t lo = 0, hi = nmesh;  // offsets[lo] <= i < offsets[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <bool COUNT, class Stack>
__device__ __forceinline__ void cast_one(const BatchMesh &bm, const LaneRay &r, float max_toi, Stack &stk,
                                         Traversal &t, uint32_t &n_wide, uint32_t &n_tri) {
    const float4 *node4 = global_float4(bm.nodes);
    const float4 *tri4 = global_float4(bm.tris);
    start_traversal(t, max_toi);
    stk.reset();
    n_wide = n_tri = 0;
    if (bm.single_leaf) {
        const float4 q0 = node4[0], q1 = node4[1];
        n_wide = 1;
        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
            visit_leaf<true>(tri4, __float_as_uint(q0.w), r, t, n_tri);
        return;
    }
    if (bm.finite && __all(ray_is_regular(r))) {
        while (!find_best_step<COUNT, 0, true>(node4, tri4, r, t, stk, n_wide, n_tri)) {
        }
    } else {
        while (!find_best_step<COUNT, 0, false>(node4, tri4, r, t, stk, n_wide, n_tri)) {
        }
    }
}

template <bool COUNT, int BLOCK, int REG>
__global__ __launch_bounds__(BLOCK) void cast_rays_batch_kernel(const BatchMesh *__restrict__ meshes,
                                                                const unsigned long long *__restrict__ offsets,
                                                                uint32_t nmesh, const pq_ray *__restrict__ rays,
                                                                uint64_t n, float max_toi,
                                                                pq_ray_hit *__restrict__ hits,
                                                                uint32_t *__restrict__ counts,
                                                                const uint32_t *__restrict__ order) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    // The binned mapping's j-th ray in sorted order; the sort keeps every
    // position inside its mesh's range, so j and i belong to the same mesh.
    const uint64_t i = order ? (uint64_t)order[j] : j;
    const uint32_t mesh = mesh_of(offsets, nmesh, j);
    const uint32_t first = __builtin_amdgcn_readfirstlane(mesh);
    const LaneRay r = load_ray(rays, i);
    LaneStack<BLOCK, REG> stk;
    stk.col = lds_stack + threadIdx.x;
    Traversal t;
    uint32_t n_wide, n_tri;
    if (__all(mesh == first)) {
        // The whole wave casts against one mesh (every wave but those that
        // straddle a mesh boundary): the record comes in by scalar loads and
        // the lane body is the single-mesh kernel's, written out in the same
        // form.  Through cast_one the compiler moved the stack push out of the
        // traversal loop and stepped a wave's lanes through the loop a few at
        // a time (3.7x the instructions, 2.5x the time on one mesh).
        const BatchMesh bm = meshes[first];
        const float4 *node4 = global_float4(bm.nodes);
        const float4 *tri4 = global_float4(bm.tris);
        stk.reset();
        start_traversal(t, max_toi);
        n_wide = n_tri = 0;
        const float4 root3 = node4[3];
        if ((__float_as_uint(root3.w) & 0x3fffffffu) == 0u) {
            const float4 q0 = node4[0], q1 = node4[1];
            n_wide = 1;
            if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
                visit_leaf<true>(tri4, __float_as_uint(q0.w), r, t, n_tri);
        } else if (bm.finite && __all(ray_is_regular(r))) {
            while (!find_best_step<COUNT, 0, true>(node4, tri4, r, t, stk, n_wide, n_tri)) {
            }
        } else {
            while (!find_best_step<COUNT, 0, false>(node4, tri4, r, t, stk, n_wide, n_tri)) {
            }
        }
        write_hit(hits, i, t, r, tri4, bm.ntri);
    } else {
        const BatchMesh bm = meshes[mesh];
        cast_one<COUNT>(bm, r, max_toi, stk, t, n_wide, n_tri);
        write_hit(hits, i, t, r, global_float4(bm.tris), bm.ntri);
    }
    if (COUNT) {
        counts[2 * i] = n_wide;
        counts[2 * i + 1] = n_tri;
    }
}

// Persistent variant: one device queue over the concatenated rays; finished
// lanes take the next rays every iteration, whatever mesh they belong to.
template <bool COUNT, int BLOCK, int REG>
__global__ __launch_bounds__(BLOCK) void cast_rays_batch_persistent_kernel(
    const BatchMesh *__restrict__ meshes, const unsigned long long *__restrict__ offsets, uint32_t nmesh,
    const pq_ray *__restrict__ rays, uint64_t n, float max_toi, pq_ray_hit *__restrict__ hits,
    uint32_t *__restrict__ counts, unsigned long long *__restrict__ queue, uint32_t chunk) {
    extern __shared__ uint32_t lds_stack[];
    LaneStack<BLOCK, REG> stk;
    stk.col = lds_stack + threadIdx.x;
    stk.reset();
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
                if (__lane_id() == 0) base = atomicAdd(&queue[0], (unsigned long long)chunk);
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
                    stk.reset();
                    n_wide = n_tri = 0;
                    active = true;
                    if (bm.single_leaf) {  // one-triangle mesh: a single test
                        const float4 *node4 = global_float4(bm.nodes);
                        const float4 q0 = node4[0], q1 = node4[1];
                        n_wide = 1;
                        if (slab_cost(q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, r, max_toi) < max_toi)
                            visit_leaf<true, 0, false>(global_float4(bm.tris), __float_as_uint(q0.w), r, t,
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
                              find_best_step<COUNT, 0, false, false>(global_float4(bm.nodes),
                                                                     global_float4(bm.tris), r, t,
                                                                     stk, n_wide, n_tri);
            if (done) {
                if (COUNT) {
                    counts[2 * ray_id] = n_wide;
                    counts[2 * ray_id + 1] = n_tri;
                }
                write_hit<false>(hits, ray_id, t, r, global_float4(bm.tris), bm.ntri);
                active = false;
            }
        }
    }
    queue_wave_done(queue, gridDim.x * (BLOCK / 64));
}

// Wave per ray over the batch: each ray's mesh by the same binary search.
template <bool COUNT>
__global__ __launch_bounds__(64) void cast_rays_batch_wave_kernel(const BatchMesh *__restrict__ meshes,
                                                                  const unsigned long long *__restrict__ offsets,
                                                                  uint32_t nmesh, const pq_ray *__restrict__ rays,
                                                                  uint64_t n, float max_toi,
                                                                  pq_ray_hit *__restrict__ hits,
                                                                  uint32_t *__restrict__ counts,
                                                                  unsigned long long *__restrict__ queue,
                                                                  uint32_t chunk) {
    uint64_t begin, end;
    while (wave_claim(queue, chunk, n, begin, end)) {
        for (uint64_t i = begin; i < end; ++i) {
            const BatchMesh bm = meshes[mesh_of(offsets, nmesh, i)];
            const LaneRay r = load_ray(rays, i);
            uint32_t n_wide, n_tri;
            const pq_ray_hit h = wave_cast_one<COUNT>(global_float4(bm.nodes), bm.num_nodes,
                                                      global_float4(bm.tris), bm.finite, bm.ntri,
                                                      bm.single_leaf != 0u, r, max_toi, n_wide, n_tri);
            if (__lane_id() == 0) {
                hits[i] = h;
                if (COUNT) {
                    counts[2 * i] = n_wide;
                    counts[2 * i + 1] = n_tri;
                }
            }
        }
    }
    queue_wave_done(queue, gridDim.x);
}

// ---------------------------------------------------------------------------
// The binned mapping's counting sort.  A ray's cell is where it enters the
// sphere inscribed in the Bvh's root box (where it passes closest to the
// sphere's centre if it misses it; its origin if it starts inside), on a 16^3
// grid over the box, numbered along a Morton curve so nearby cells get nearby
// numbers.  The cell only orders the work: every ray still runs its own
// find_best, so the hits are those of the static mapping.  The rays are split
// into `blocks` contiguous ranges of `per_block` rays, one sort workgroup each,
// and no global atomics are used:
//   bin_count_kernel   per range: each ray's cell and its rank among the
//                      range's rays of that cell (LDS atomics), and the
//                      range's count per cell into hist[range * kRayBins + cell];
//   bin_scan_kernel    per 64 cells (one workgroup, a cell per lane, the ranges
//                      split over its waves): exclusive prefix sum over the
//                      ranges in place, and each cell's total;
//   bin_cells_kernel   exclusive prefix sum of the totals: each cell's first
//                      sorted position;
//   bin_scatter_kernel order[first[cell] + hist[range * kRayBins + cell] + rank]
//                      = ray index.
// A batch sorts per mesh: ranges of kRayBinRange rays of one mesh each
// (bin_count_batch_kernel, cells of that mesh's root box), the scan over each
// mesh's ranges (a grid row per mesh), each mesh's cell offsets starting at its
// first ray (one bin_cells_kernel workgroup per mesh), and the scatter finding
// a ray's mesh and range from its index (bin_scatter_batch_kernel).
// ---------------------------------------------------------------------------
constexpr int kBinBlock = 256;          // scatter workgroup
constexpr int kCountBlock = 1024;       // count workgroup
constexpr int kCountPer = 4;            // rays per count lane per round (loads in flight together)
constexpr uint32_t kBinRange = kRayBinRange;  // rays per sort range (before the block cap)
static_assert(kBinRange == kCountBlock * kCountPer, "one round of loads covers a range");

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bit k -> bit 3k (4 bits)
    v &= 0xfu;
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}

__device__ __forceinline__ uint32_t ray_cell(const float mn[3], const float mx[3], const LaneRay &r) {
    float c[3], rad = FLT_MAX;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5f * (mn[a] + mx[a]);
        rad = fminf(rad, 0.5f * (mx[a] - mn[a]));
    }
    const float oc[3] = {r.ox - c[0], r.oy - c[1], r.oz - c[2]};
    const float dd = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    const float b = oc[0] * r.dx + oc[1] * r.dy + oc[2] * r.dz;
    const float cc = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - rad * rad;
    float t = dd > 0.0f ? (-b - sqrtf(fmaxf(b * b - dd * cc, 0.0f))) / dd : 0.0f;
    t = fmaxf(t, 0.0f);  // NaN -> 0
    const float p[3] = {r.ox + t * r.dx, r.oy + t * r.dy, r.oz + t * r.dz};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ext = mx[a] - mn[a];
        const float u = ext > 0.0f ? (p[a] - mn[a]) / ext * (float)(1u << kRayBinBits) : 0.0f;
        q[a] = (uint32_t)fminf(fmaxf(u, 0.0f), (float)((1u << kRayBinBits) - 1u));
    }
    return (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
}

// Ranks rays [begin, end) in the cells of the sphere inscribed in the root box
// of node4's tree: keyrank[j] = cell | rank << kRayBinKeyBits, counts in cnt.
__device__ __forceinline__ void count_cells(const float4 *node4, const pq_ray *__restrict__ rays, uint64_t begin,
                                            uint64_t end, uint32_t *cnt, uint32_t *__restrict__ keyrank) {
    const float4 q0 = node4[0], q1 = node4[1], q2 = node4[2], q3 = node4[3];
    const bool two = (__float_as_uint(q3.w) & 0x3fffffffu) != 0u;  // a root with one leaf has no right child
    const float mn[3] = {two ? fminf(q0.x, q2.x) : q0.x, two ? fminf(q0.y, q2.y) : q0.y,
                         two ? fminf(q0.z, q2.z) : q0.z};
    const float mx[3] = {two ? fmaxf(q1.x, q3.x) : q1.x, two ? fmaxf(q1.y, q3.y) : q1.y,
                         two ? fmaxf(q1.z, q3.z) : q1.z};
    for (uint64_t base = begin; base < end; base += kBinRange) {
        RawRay raw[kCountPer];
#pragma unroll
        for (int k = 0; k < kCountPer; ++k) {
            const uint64_t j = base + threadIdx.x + (uint64_t)k * kCountBlock;
            if (j < end) raw[k] = load_raw_ray(rays, j);
        }
#pragma unroll
        for (int k = 0; k < kCountPer; ++k) {
            const uint64_t j = base + threadIdx.x + (uint64_t)k * kCountBlock;
            if (j < end) {
                const uint32_t cell = ray_cell(mn, mx, lane_ray(raw[k]));
                const uint32_t rank = atomicAdd(&cnt[cell], 1u);
                keyrank[j] = cell | (rank << kRayBinKeyBits);
            }
        }
    }
}

__global__ __launch_bounds__(kCountBlock) void bin_count_kernel(const pq_bvh_node_wide *__restrict__ nodes,
                                                                const pq_ray *__restrict__ rays, uint64_t n,
                                                                uint64_t per_block, uint32_t *__restrict__ hist,
                                                                uint32_t *__restrict__ keyrank) {
    __shared__ uint32_t cnt[kRayBins];
    for (uint32_t c = threadIdx.x; c < kRayBins; c += kCountBlock) cnt[c] = 0u;
    __syncthreads();
    const uint64_t begin = (uint64_t)blockIdx.x * per_block;
    const uint64_t end = begin + per_block < n ? begin + per_block : n;
    count_cells(reinterpret_cast<const float4 *>(nodes), rays, begin, end, cnt, keyrank);
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < kRayBins; c += kCountBlock) hist[(uint64_t)blockIdx.x * kRayBins + c] = cnt[c];
}

// A batch: one workgroup per sort range; range r holds rays of mesh
// range_mesh[r] only, and the cells are those of that mesh's root box.
__global__ __launch_bounds__(kCountBlock) void bin_count_batch_kernel(
    const BatchMesh *__restrict__ meshes, const unsigned long long *__restrict__ offsets,
    const uint32_t *__restrict__ range_mesh, const uint32_t *__restrict__ range_first,
    const pq_ray *__restrict__ rays, uint32_t *__restrict__ hist, uint32_t *__restrict__ keyrank) {
    __shared__ uint32_t cnt[kRayBins];
    for (uint32_t c = threadIdx.x; c < kRayBins; c += kCountBlock) cnt[c] = 0u;
    __syncthreads();
    const uint32_t m = range_mesh[blockIdx.x];
    const uint64_t begin = offsets[m] + (uint64_t)(blockIdx.x - range_first[m]) * kBinRange;
    const uint64_t end = begin + kBinRange < offsets[m + 1] ? begin + kBinRange : offsets[m + 1];
    count_cells(global_float4(meshes[m].nodes), rays, begin, end, cnt, keyrank);
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < kRayBins; c += kCountBlock) hist[(uint64_t)blockIdx.x * kRayBins + c] = cnt[c];
}

// 64 cells per workgroup, one per lane; its kScanWaves waves split the ranges
// into consecutive chunks.  Each wave sums its chunk of a cell's counts (one
// coalesced 256-byte row of 64 cells per range), the chunk totals are
// combined in LDS, and each wave then writes its chunk's exclusive prefixes
// in place.  The last wave's running total is the cell's total.  With
// range_first (a batch), blockIdx.y is the mesh: its ranges are
// [range_first[m], range_first[m + 1]) and its totals start at m * kRayBins.
constexpr uint32_t kScanWaves = 16;
__global__ __launch_bounds__(64 * kScanWaves) void bin_scan_kernel(uint32_t *__restrict__ hist, uint32_t blocks,
                                                                  const uint32_t *__restrict__ range_first,
                                                                  uint32_t *__restrict__ totals) {
    __shared__ uint32_t part[kScanWaves][64];
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    const uint32_t cell = blockIdx.x * 64u + lane;
    uint32_t lo = 0, span = blocks;
    if (range_first) {
        lo = range_first[blockIdx.y];
        span = range_first[blockIdx.y + 1] - lo;
        totals += (uint64_t)blockIdx.y * kRayBins;
    }
    const uint32_t chunk = (span + kScanWaves - 1u) / kScanWaves;
    const uint32_t r0 = lo + (wave * chunk < span ? wave * chunk : span);
    const uint32_t r1 = lo + (wave * chunk + chunk < span ? wave * chunk + chunk : span);
    uint32_t sum = 0;
    for (uint32_t base = r0; base < r1; base += 16u) {
        uint32_t v[16];
#pragma unroll
        for (uint32_t k = 0; k < 16u; ++k) v[k] = base + k < r1 ? hist[(uint64_t)(base + k) * kRayBins + cell] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 16u; ++k) sum += v[k];
    }
    part[wave][lane] = sum;
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t w = 0; w < wave; ++w) carry += part[w][lane];
    for (uint32_t base = r0; base < r1; base += 16u) {
        uint32_t v[16];
#pragma unroll
        for (uint32_t k = 0; k < 16u; ++k) v[k] = base + k < r1 ? hist[(uint64_t)(base + k) * kRayBins + cell] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 16u; ++k) {
            if (base + k < r1) {
                hist[(uint64_t)(base + k) * kRayBins + cell] = carry;
                carry += v[k];
            }
        }
    }
    if (wave == kScanWaves - 1u) totals[cell] = carry;
}

// One workgroup: totals -> each cell's first sorted position, in place.  With
// offsets (a batch), workgroup m does mesh m's cells, starting at its first ray.
__global__ __launch_bounds__(kBinBlock) void bin_cells_kernel(uint32_t *__restrict__ totals,
                                                              const unsigned long long *__restrict__ offsets) {
    static_assert(kRayBins % kBinBlock == 0, "cells split evenly over the workgroup");
    constexpr uint32_t kPer = kRayBins / kBinBlock;
    __shared__ uint32_t wave_sum[kBinBlock / 64];
    const uint32_t t = threadIdx.x, lane = t % 64u, wave = t / 64u;
    totals += (uint64_t)blockIdx.x * kRayBins;
    uint32_t c[kPer], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        c[k] = totals[kPer * t + k];
        s += c[k];
    }
    uint32_t incl = s;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63u) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t run = (offsets ? (uint32_t)offsets[blockIdx.x] : 0u) + incl - s;
    for (uint32_t w = 0; w < wave; ++w) run += wave_sum[w];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        totals[kPer * t + k] = run;
        run += c[k];
    }
}

__global__ __launch_bounds__(kBinBlock) void bin_scatter_kernel(const uint32_t *__restrict__ keyrank,
                                                                const uint32_t *__restrict__ hist,
                                                                const uint32_t *__restrict__ first, uint64_t n,
                                                                uint64_t per_block, uint32_t blocks,
                                                                uint32_t *__restrict__ order) {
    const uint64_t j = (uint64_t)blockIdx.x * kBinBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t kr = keyrank[j];
    const uint32_t cell = kr & (kRayBins - 1u), rank = kr >> kRayBinKeyBits;
    order[first[cell] + hist[(j / per_block) * kRayBins + cell] + rank] = (uint32_t)j;
}

__global__ __launch_bounds__(kBinBlock) void bin_scatter_batch_kernel(
    const uint32_t *__restrict__ keyrank, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ first,
    const unsigned long long *__restrict__ offsets, uint32_t nmesh, const uint32_t *__restrict__ range_first,
    uint64_t n, uint32_t *__restrict__ order) {
    const uint64_t j = (uint64_t)blockIdx.x * kBinBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t kr = keyrank[j];
    const uint32_t cell = kr & (kRayBins - 1u), rank = kr >> kRayBinKeyBits;
    const uint32_t m = mesh_of(offsets, nmesh, j);
    const uint64_t range = range_first[m] + (j - offsets[m]) / kBinRange;
    order[first[(uint64_t)m * kRayBins + cell] + hist[range * kRayBins + cell] + rank] = (uint32_t)j;
}

// The binned mapping's sort: leaves in *order the ray index at each sorted position.
hipError_t launch_ray_bins(const DeviceMesh &m, const pq_ray *rays, uint64_t n, hipStream_t stream,
                           const uint32_t **order) {
    uint64_t blocks = (n + kBinRange - 1) / kBinRange;
    if (blocks > kRayBinMaxBlocks) blocks = kRayBinMaxBlocks;
    const uint64_t per_block = (n + blocks - 1) / blocks;
    RayBinSpace ws{};
    hipError_t e = ray_bin_workspace(stream, n, (uint32_t)blocks, &ws);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bin_count_kernel, dim3((uint32_t)blocks), dim3(kCountBlock), 0, stream, m.nodes, rays, n,
                       per_block, ws.hist, ws.keyrank);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(kRayBins / 64), dim3(64 * kScanWaves), 0, stream, ws.hist, (uint32_t)blocks,
                       nullptr, ws.totals);
    hipLaunchKernelGGL(bin_cells_kernel, dim3(1), dim3(kBinBlock), 0, stream, ws.totals, nullptr);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3((uint32_t)((n + kBinBlock - 1) / kBinBlock)), dim3(kBinBlock), 0,
                       stream, ws.keyrank, ws.hist, ws.totals, n, per_block, (uint32_t)blocks, ws.order);
    *order = ws.order;
    return hipGetLastError();
}

// A batch's sort, per mesh, into the batch's own buffers (b.bins).
hipError_t launch_batch_bins(const DeviceBatch &b, const pq_ray *rays, hipStream_t stream) {
    const BatchBins &bb = b.bins;
    if (!bb.order) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bin_count_batch_kernel, dim3(bb.num_ranges), dim3(kCountBlock), 0, stream, b.meshes, b.offsets,
                       bb.range_mesh, bb.range_first, rays, bb.hist, bb.keyrank);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(kRayBins / 64, b.num_meshes), dim3(64 * kScanWaves), 0, stream, bb.hist,
                       0u, bb.range_first, bb.first);
    hipLaunchKernelGGL(bin_cells_kernel, dim3(b.num_meshes), dim3(kBinBlock), 0, stream, bb.first, b.offsets);
    hipLaunchKernelGGL(bin_scatter_batch_kernel, dim3((uint32_t)((b.num_rays + kBinBlock - 1) / kBinBlock)),
                       dim3(kBinBlock), 0, stream, bb.keyrank, bb.hist, bb.first, b.offsets, b.num_meshes,
                       bb.range_first, b.num_rays, bb.order);
    return hipGetLastError();
}

// Calls f(std::integral_constant<int, BLOCK>, std::integral_constant<int, REG>)
// for a tuning's block size (64, 128, 256) and register stack depth (0, 1, 2, 4).
template <class F>
hipError_t with_shape(const pq_ray_tuning &tu, F &&f) {
    using std::integral_constant;
    const uint32_t block = tu.block_size ? tu.block_size : 256u;
    switch (block * 8u + tu.stack_regs) {
#define PQ_SHAPE_CASE(B, R) \
    case B * 8u + R: return f(integral_constant<int, B>{}, integral_constant<int, R>{});
        PQ_SHAPE_CASE(64, 0) PQ_SHAPE_CASE(64, 1) PQ_SHAPE_CASE(64, 2) PQ_SHAPE_CASE(64, 4)
        PQ_SHAPE_CASE(128, 0) PQ_SHAPE_CASE(128, 1) PQ_SHAPE_CASE(128, 2) PQ_SHAPE_CASE(128, 4)
        PQ_SHAPE_CASE(256, 0) PQ_SHAPE_CASE(256, 1) PQ_SHAPE_CASE(256, 2) PQ_SHAPE_CASE(256, 4)
#undef PQ_SHAPE_CASE
        default: return hipErrorInvalidValue;
    }
}

// LDS bytes of a workgroup's stacks: depth-1 entries per lane, less the REG in registers.
size_t stack_lds_bytes(uint32_t depth, int block, int regs) {
    const uint32_t entries = ray_stack_entries(depth);
    const uint32_t in_lds = entries > (uint32_t)regs ? entries - (uint32_t)regs : 1u;
    return (size_t)in_lds * (size_t)block * sizeof(uint32_t);
}

// Workgroups for a persistent grid: the occupancy limit (or the tuning's cap)
// per CU, and no more waves than chunks of work.
hipError_t persistent_blocks(const void *kernel, int block, size_t lds, const pq_ray_tuning &tu, uint64_t work_units,
                             uint64_t *blocks) {
    hipError_t e;
    int dev = 0, cus = 0, per_cu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds)) != hipSuccess) return e;
    if (tu.blocks_per_cu > 0 && (int)tu.blocks_per_cu < per_cu) per_cu = (int)tu.blocks_per_cu;
    if (per_cu < 1) per_cu = 1;
    const uint64_t waves_per_block = (uint64_t)block / 64u;
    uint64_t b = (uint64_t)cus * (uint64_t)per_cu;
    if (b * waves_per_block > work_units) b = (work_units + waves_per_block - 1) / waves_per_block;
    *blocks = b < 1 ? 1 : b;
    return hipSuccess;
}

}  // namespace

uint32_t ray_stack_entries(uint32_t depth) { return depth > 1 ? depth - 1 : 1; }

// The wave mapping's stack is one VGPR (64 entries).
bool wave_mapping_fits(uint32_t depth) { return ray_stack_entries(depth) <= 64u; }

hipError_t launch_cast_rays_culling(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi,
                                    uint32_t culling, pq_ray_hit *hits, uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    constexpr int B = 256;
    const size_t lds = stack_lds_bytes(m.depth, B, 0);
    const dim3 grid((uint32_t)((n + B - 1) / B)), block(B);
    // the mesh's binned mapping sorts these rays too; any other mapping runs static lanes
    const bool single_leaf = m.num_nodes == 1 && m.num_triangles == 1;
    const uint32_t *order = nullptr;
    if (m.tuning.mapping == PQ_RAY_MAPPING_BINNED && !single_leaf && n <= 0xffffffffull) {
        const hipError_t e = launch_ray_bins(m, rays, n, stream, &order);
        if (e != hipSuccess) return e;
    }
    if (culling == 1 && counts)
        hipLaunchKernelGGL((cast_rays_kernel<true, 1, B, 0>), grid, block, lds, stream, m.nodes, m.tris,
                           m.num_triangles, rays, n, max_toi, hits, counts, order, m.finite);
    else if (culling == 1)
        hipLaunchKernelGGL((cast_rays_kernel<false, 1, B, 0>), grid, block, lds, stream, m.nodes, m.tris,
                           m.num_triangles, rays, n, max_toi, hits, counts, order, m.finite);
    else if (counts)
        hipLaunchKernelGGL((cast_rays_kernel<true, 2, B, 0>), grid, block, lds, stream, m.nodes, m.tris,
                           m.num_triangles, rays, n, max_toi, hits, counts, order, m.finite);
    else
        hipLaunchKernelGGL((cast_rays_kernel<false, 2, B, 0>), grid, block, lds, stream, m.nodes, m.tris,
                           m.num_triangles, rays, n, max_toi, hits, counts, order, m.finite);
    return hipGetLastError();
}

namespace {
template <bool COUNT>
hipError_t launch_single(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi, pq_ray_hit *hits,
                         uint32_t *counts, hipStream_t stream) {
    const pq_ray_tuning &tu = m.tuning;
    const bool single_leaf = m.num_nodes == 1 && m.num_triangles == 1;
    uint32_t mapping = tu.mapping;
    if (mapping == PQ_RAY_MAPPING_WAVE && !wave_mapping_fits(m.depth)) mapping = PQ_RAY_MAPPING_PERSISTENT;
    if (mapping == PQ_RAY_MAPPING_PERSISTENT && single_leaf) mapping = PQ_RAY_MAPPING_STATIC;
    if (mapping == PQ_RAY_MAPPING_BINNED && (single_leaf || n > 0xffffffffull)) mapping = PQ_RAY_MAPPING_STATIC;
    if (mapping == PQ_RAY_MAPPING_WAVE) {
        unsigned long long *queue = nullptr;
        hipError_t e = stream_queue(stream, &queue);
        if (e != hipSuccess) return e;
        const uint32_t chunk = tu.chunk ? tu.chunk : 4u;
        uint64_t blocks = 0;
        if ((e = persistent_blocks((const void *)cast_rays_wave_kernel<COUNT>, 64, 0, tu, (n + chunk - 1) / chunk,
                                   &blocks)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(cast_rays_wave_kernel<COUNT>, dim3((uint32_t)blocks), dim3(64), 0, stream, m.nodes,
                           m.num_nodes, m.tris, m.num_triangles, rays, n, max_toi, hits, counts, queue, chunk,
                           m.finite);
        return hipGetLastError();
    }
    return with_shape(tu, [&](auto bc, auto rc) -> hipError_t {
        constexpr int B = decltype(bc)::value, R = decltype(rc)::value;
        const size_t lds = stack_lds_bytes(m.depth, B, R);
        if (mapping == PQ_RAY_MAPPING_PERSISTENT) {
            unsigned long long *queue = nullptr;
            hipError_t e = stream_queue(stream, &queue);
            if (e != hipSuccess) return e;
            const uint32_t chunk = tu.chunk ? tu.chunk : 128u;
            uint64_t blocks = 0;
            if ((e = persistent_blocks((const void *)cast_rays_persistent_kernel<COUNT, B, R>, B, lds, tu,
                                       (n + chunk - 1) / chunk, &blocks)) != hipSuccess)
                return e;
            hipLaunchKernelGGL((cast_rays_persistent_kernel<COUNT, B, R>), dim3((uint32_t)blocks), dim3(B), lds,
                               stream, m.nodes, m.tris, m.num_triangles, rays, n, max_toi, hits, counts, queue, chunk);
        } else {
            const uint32_t *order = nullptr;
            if (mapping == PQ_RAY_MAPPING_BINNED) {
                const hipError_t e = launch_ray_bins(m, rays, n, stream, &order);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL((cast_rays_kernel<COUNT, 0, B, R>), dim3((uint32_t)((n + B - 1) / B)), dim3(B), lds,
                               stream, m.nodes, m.tris, m.num_triangles, rays, n, max_toi, hits, counts, order,
                               m.finite);
        }
        return hipGetLastError();
    });
}

template <bool COUNT>
hipError_t launch_batch(const DeviceBatch &b, const pq_ray *rays, float max_toi, pq_ray_hit *hits, uint32_t *counts,
                        hipStream_t stream) {
    const uint64_t n = b.num_rays;
    const pq_ray_tuning &tu = b.tuning;
    uint32_t mapping = tu.mapping;
    if (mapping == PQ_RAY_MAPPING_WAVE && !wave_mapping_fits(b.max_depth)) mapping = PQ_RAY_MAPPING_PERSISTENT;
    if (mapping == PQ_RAY_MAPPING_WAVE) {
        const uint32_t chunk = tu.chunk ? tu.chunk : 4u;
        uint64_t blocks = 0;
        hipError_t e;
        if ((e = persistent_blocks((const void *)cast_rays_batch_wave_kernel<COUNT>, 64, 0, tu,
                                   (n + chunk - 1) / chunk, &blocks)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(cast_rays_batch_wave_kernel<COUNT>, dim3((uint32_t)blocks), dim3(64), 0, stream, b.meshes,
                           b.offsets, b.num_meshes, rays, n, max_toi, hits, counts, b.queue, chunk);
        return hipGetLastError();
    }
    const uint32_t *order = nullptr;
    if (mapping == PQ_RAY_MAPPING_BINNED) {
        hipError_t e = launch_batch_bins(b, rays, stream);
        if (e != hipSuccess) return e;
        order = b.bins.order;
    }
    return with_shape(tu, [&](auto bc, auto rc) -> hipError_t {
        constexpr int B = decltype(bc)::value, R = decltype(rc)::value;
        const size_t lds = stack_lds_bytes(b.max_depth, B, R);
        if (mapping == PQ_RAY_MAPPING_PERSISTENT) {
            const uint32_t chunk = tu.chunk ? tu.chunk : 128u;
            uint64_t blocks = 0;
            hipError_t e;
            if ((e = persistent_blocks((const void *)cast_rays_batch_persistent_kernel<COUNT, B, R>, B, lds, tu,
                                       (n + chunk - 1) / chunk, &blocks)) != hipSuccess)
                return e;
            hipLaunchKernelGGL((cast_rays_batch_persistent_kernel<COUNT, B, R>), dim3((uint32_t)blocks), dim3(B), lds,
                               stream, b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts, b.queue,
                               chunk);
        } else {
            hipLaunchKernelGGL((cast_rays_batch_kernel<COUNT, B, R>), dim3((uint32_t)((n + B - 1) / B)), dim3(B), lds,
                               stream, b.meshes, b.offsets, b.num_meshes, rays, n, max_toi, hits, counts, order);
        }
        return hipGetLastError();
    });
}
}  // namespace

hipError_t launch_cast_rays(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi, pq_ray_hit *hits,
                            uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    return counts ? launch_single<true>(m, rays, n, max_toi, hits, counts, stream)
                  : launch_single<false>(m, rays, n, max_toi, hits, counts, stream);
}

hipError_t launch_cast_rays_batch(const DeviceBatch &b, const pq_ray *rays, float max_toi, pq_ray_hit *hits,
                                  uint32_t *counts, hipStream_t stream) {
    if (b.num_rays == 0) return hipSuccess;
    return counts ? launch_batch<true>(b, rays, max_toi, hits, counts, stream)
                  : launch_batch<false>(b, rays, max_toi, hits, counts, stream);
}

}  // namespace pq
