// contact.hip -- batched query::contact on shape pairs (gfx950).
//
// Stage 1 (one lane per pair): pos12 = pos1.inv_mul(pos2), the dispatch of
// DefaultQueryDispatcher::contact (src/query/default_query_dispatcher.rs:302-356)
// for the shapes held here, then
//   * Ball-Ball: the closed form of contact_ball_ball.rs:9-42;
//   * Ball with a Cuboid, Capsule or Cone: the closed-form projection of the
//     ball centre (point_aabb.rs, point_capsule.rs, point_cone.rs) and
//     contact_ball_convex_polyhedron.rs:12-63; with a ConvexPolyhedron, GJK
//     of the shape against ConstantOrigin (point_support_map.rs:17-51), the
//     centre inside going to the EPA queue like an intersecting pair;
//   * support-map pairs: contact_support_map_support_map_with_params
//     (src/query/contact/contact_support_map_support_map.rs:40-77) -- seed
//     direction (caller-supplied or parry's normalize(pos12.t) / +x), then GJK
//     in registers.  Pairs GJK finds intersecting are appended, with their
//     final simplex, to an EPA queue (one atomic per wave: ballot + masked
//     popcount give each lane its slot).
// Stage 2 (one wave per queued pair): EPA (src/query/epa/epa3.rs:428-651) with
//   the polytope -- vertices, faces, adjacency, the face heap (the same binary
//   heap algorithm as Rust's BinaryHeap, so equal-distance faces pop in parry's
//   order) and the silhouette stack -- in LDS.  The sequential algorithm runs
//   identically on all 64 lanes; the ConvexPolyhedron support scans are split
//   across the lanes and merged with a wave reduction that keeps parry's
//   first-strict-maximum rule.  A pair whose polytope outgrows the LDS
//   capacity is re-run by stage 3, the same code with large buffers in global
//   memory.
// Every contact is finished as contact_support_map_support_map (dist, point2,
// normal2 in shape 2's frame) and Contact::transform_by_mut (contact.rs:171-176).
#include <cstring>

#include "contact_gjk.h"

namespace pq {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kUnsupportedPair = 0xffffffffu;

struct DevContact {
    float p1[3], p2[3], n1[3], n2[3], dist;
    uint32_t has;
};
static_assert(sizeof(DevContact) == sizeof(pq_contact), "pq_contact layout");

// Queued EPA work item: pair index, what the EPA solves and GJK's final simplex.
enum : uint32_t {
    kPairSupportMaps = 0,  // contact_support_map_support_map on (shape 1, shape 2)
    kBallSecond = 1,       // (ConvexPolyhedron, Ball): project the ball centre
    kBallFirst = 2,        // (Ball, ConvexPolyhedron): the same, flipped
    kPairSupportMapsLocal = 3,  // contact_support_map_support_map on any pair, Ball included,
                                // results left in the shapes' local frames (the shape cast's
                                // impact geometry, shape_cast_support_map_support_map.rs:42-46)
    kPairFeatureMaps = 4,       // the same on the pair's polygonal feature maps (a Capsule is its
                                // segment), for contact manifolds (contact_manifolds_pfm_pfm.rs:66-74)
};
__device__ __forceinline__ bool is_ball_mode(uint32_t mode) { return mode == kBallFirst || mode == kBallSecond; }
__device__ __forceinline__ bool is_local_mode(uint32_t mode) {
    return mode == kPairSupportMapsLocal || mode == kPairFeatureMaps;
}
// A shape's polygonal feature map and border radius (shape.rs as_polygonal_feature_map):
// Cuboid, Cone and Triangle are their own, a Capsule is its segment with the radius as
// border, a ConvexPolyhedron is its own when it has faces (its face count, as uint32
// bits in param[1], is not zero).  False for shapes without one here (Ball).
__device__ __forceinline__ bool feature_map(ShapeRef &g, float &border) {
    border = 0.0f;
    if (g.type == PQ_SHAPE_CUBOID || g.type == PQ_SHAPE_CONE || g.type == PQ_SHAPE_TRIANGLE) return true;
    if (g.type == PQ_SHAPE_CONVEX_POLYHEDRON) return __float_as_uint(g.p1) != 0u;
    if (g.type == PQ_SHAPE_CAPSULE) {
        g.type = kShapeSegment;
        border = g.p1;
        return true;
    }
    return false;
}
struct EpaItem {
    uint32_t pair;
    uint32_t dim;
    uint32_t mode;
    uint32_t pad;
    float v[4][9];  // point, orig1, orig2
};

// Writes a contact record straight to its output slot (a local record copied
// out afterwards would be kept in scratch memory).
__device__ __forceinline__ void store_contact(DevContact *c, f3 w1, f3 w2, f3 m1, f3 m2, float dist) {
    c->p1[0] = w1.x; c->p1[1] = w1.y; c->p1[2] = w1.z;
    c->p2[0] = w2.x; c->p2[1] = w2.y; c->p2[2] = w2.z;
    c->n1[0] = m1.x; c->n1[1] = m1.y; c->n1[2] = m1.z;
    c->n2[0] = m2.x; c->n2[1] = m2.y; c->n2[2] = m2.z;
    c->dist = dist;
    c->has = 1;
}

__device__ __forceinline__ DevContact no_contact(uint32_t has) {
    DevContact c;
    for (int k = 0; k < 3; ++k) c.p1[k] = c.p2[k] = c.n1[k] = c.n2[k] = 0.0f;
    c.dist = 0.0f;
    c.has = has;
    return c;
}

__device__ __forceinline__ void finish_contact(const iso &pos1, const iso &pos2, const iso &pos12, f3 p1, f3 p2_1,
                                               f3 n1, bool local, DevContact *o) {
    // contact_support_map_support_map (contact_support_map_support_map.rs:18-26)
    const float dist = dot3(sub3(p2_1, p1), n1);
    const f3 p2 = qrot_inv(pos12.q, sub3(p2_1, pos12.t));
    const f3 n2 = qrot_inv(pos12.q, neg3(n1));
    if (local) {
        store_contact(o, p1, p2, n1, n2, dist);
        return;
    }
    // Contact::transform_by_mut
    store_contact(o, iso_point(pos1, p1), iso_point(pos2, p2), qrot(pos1.q, n1), qrot(pos2.q, n2), dist);
}

// Builds the Ball-vs-convex contact (contact_ball_convex_polyhedron.rs:27-63)
// and maps it to world space; pb = pose of the ball in the convex shape's frame.
__device__ __forceinline__ void finish_ball_contact(const iso &pos1, const iso &pos2, const iso &pb, float radius,
                                                    float prediction, const PointProj &proj, bool has_fn, f3 fnrm,
                                                    uint32_t mode, DevContact *o) {
    f3 p1, p2, n1, n2;
    float dist;
    if (!convex_ball_contact(pb, radius, prediction, proj, has_fn, fnrm, p1, p2, n1, n2, dist)) {
        *o = no_contact(0);
        return;
    }
    const bool flip = mode == kBallFirst;  // Contact::flipped (contact.rs:156-166)
    const f3 q1 = sel2(flip, p2, p1), q2 = sel2(flip, p1, p2), m1 = sel2(flip, n2, n1), m2 = sel2(flip, n1, n2);
    store_contact(o, iso_point(pos1, q1), iso_point(pos2, q2), qrot(pos1.q, m1), qrot(pos2.q, m2), dist);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// counters[0..3]: EPA queue length and the three overflow-list lengths;
// counters[4..7]: the next list entry each EPA tier hands out;
// counters[8]: continuation records taken by the first tier.
constexpr uint32_t kNumCounters = 9;
__global__ void reset_counters_kernel(uint32_t *counters) {
    if (threadIdx.x < kNumCounters) counters[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Stage 1
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void contact_gjk_kernel(
    const ShapeRecord *__restrict__ rec, const float4 *__restrict__ verts, const float *__restrict__ poses,
    const uint32_t *__restrict__ pairs, uint64_t m, const unsigned long long *__restrict__ m_dev, float prediction,
    const float *__restrict__ seeds, DevContact *__restrict__ out, uint32_t *__restrict__ iters,
    EpaItem *__restrict__ queue, uint32_t *__restrict__ counters, uint32_t flags) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m_dev) {
        const uint64_t md = *m_dev;
        if (md < m) m = md;
    }
    bool need_epa = false;
    uint32_t ball_mode = kPairSupportMaps;
    Simplex s;
    uint32_t pair_id = 0;
    if (k < m) {
        pair_id = (uint32_t)k;
        const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
        const iso pos12 = iso_inv_mul(pos1, pos2);
        ShapeRef g1 = load_shape(rec, verts, i), g2 = load_shape(rec, verts, j);
        DevContact *const c = out + k;
        uint32_t gi = 0;
        const bool pfm = (flags & kContactFeatureMaps) != 0u;
        bool unsupported = false;
        float total_prediction = prediction;
        if (pfm) {  // total_prediction = prediction + border1 + border2 (contact_manifolds_pfm_pfm.rs:65)
            float br1 = 0.0f, br2 = 0.0f;
            unsupported = !feature_map(g1, br1) || !feature_map(g2, br2);
            total_prediction = prediction + br1 + br2;
        }
        // What GJK runs on, if anything: a support-map pair, or a ConvexPolyhedron
        // against ConstantOrigin (the ball-centre projection).  One call site below.
        bool run_gjk = false;
        iso gjk_pos{}, seed_pos{};
        ShapeRef ga{}, gb{};
        float max_dist = prediction;
        f3 dir{1.0f, 0.0f, 0.0f}, dflt{1.0f, 0.0f, 0.0f};
        bool seeded = false;
        // Ball pairs: the pose of the ball in the convex shape's frame, the shape, the radius.
        iso pb{};
        ShapeRef cv{};
        float radius = 0.0f;
        const bool local = (flags & (kContactLocalSupportMaps | kContactFeatureMaps)) != 0u;
        if (unsupported) {
            *c = no_contact(kUnsupportedPair);
        } else if (!local && g1.type == PQ_SHAPE_BALL && g2.type == PQ_SHAPE_BALL) {
            // contact_ball_ball
            const float r1 = g1.p0, r2 = g2.p0;
            const f3 c21 = pos12.t;
            const float d2 = norm_sq3(c21);
            const float sum = r1 + r2;
            const float sum_err = sum + prediction;
            if (d2 < sum_err * sum_err) {
                const f3 n1 = d2 != 0.0f ? div3(c21, sqrtf(norm_sq3(c21))) : f3{1.0f, 0.0f, 0.0f};
                const f3 n2 = neg3(qrot_inv(pos12.q, n1));
                const f3 a = mul3(n1, r1), b = mul3(n2, r2);
                store_contact(c, iso_point(pos1, a), iso_point(pos2, b), qrot(pos1.q, n1), qrot(pos2.q, n2),
                              sqrtf(d2) - sum);
            } else {
                *c = no_contact(0);
            }
        } else if (!local && (g1.type == PQ_SHAPE_BALL || g2.type == PQ_SHAPE_BALL)) {
            // contact_ball_convex_polyhedron / contact_convex_polyhedron_ball
            // (default_query_dispatcher.rs:326-335): a Ball first is the flipped
            // contact of (shape 2, ball) under pos12^-1.
            if (g1.type == PQ_SHAPE_BALL) {
                ball_mode = kBallFirst;
                pb = iso_inverse(pos12);
                cv = g2;
                radius = g1.p0;
            } else {
                ball_mode = kBallSecond;
                pb = pos12;
                cv = g1;
                radius = g2.p0;
            }
            const f3 ctr = pb.t;
            PointProj proj{false, f3{0.0f, 0.0f, 0.0f}};
            uint32_t fk = CF_UNKNOWN, fid = 0;
            if (cv.type == PQ_SHAPE_CONVEX_POLYHEDRON) {
                // local_point_projection_on_support_map (point_support_map.rs:17-51)
                run_gjk = true;
                ga = cv;
                gb = origin_shape();
                seed_pos = point_proj_seed_iso(ctr);
                gjk_pos = point_proj_iso(ctr);
                max_dist = kFltMax;
                float l;
                if (!try_unit(ctr, dir, l)) dir = f3{1.0f, 0.0f, 0.0f};
            } else {
                bool has_fn = false;
                f3 fnrm{0.0f, 0.0f, 0.0f};
                if (cv.type == PQ_SHAPE_CUBOID) {
                    proj = cuboid_point_feature(cv, ctr, fk, fid);
                    has_fn = cuboid_feature_normal(fk, fid, fnrm);
                } else if (cv.type == PQ_SHAPE_TRIANGLE) {
                    proj = triangle_point(cv, ctr);
                    has_fn = triangle_normal(cv, fnrm);  // any feature (shape.rs:919-928)
                } else if (cv.type == PQ_SHAPE_CAPSULE) {
                    proj = capsule_point(cv, ctr);
                } else {
                    proj = cone_point(cv, ctr);
                }
                finish_ball_contact(pos1, pos2, pb, radius, prediction, proj, has_fn, fnrm, ball_mode, c);
            }
        } else {
            run_gjk = true;
            ball_mode = pfm ? kPairFeatureMaps : (local ? kPairSupportMapsLocal : kPairSupportMaps);
            max_dist = total_prediction;
            ga = g1;
            gb = g2;
            seed_pos = gjk_pos = pos12;
            // parry's default seed: normalize(pos12.t), else +x (contact_support_map_support_map.rs:52-60)
            const float tsq = norm_sq3(pos12.t);
            dflt = tsq > kEps * kEps ? div3(pos12.t, sqrtf(tsq)) : f3{1.0f, 0.0f, 0.0f};
            dir = dflt;
            if (seeds) {
                const f3 sd = f3{seeds[3 * k], seeds[3 * k + 1], seeds[3 * k + 2]};
                seeded = norm_sq3(sd) > kEps * kEps;  // a zero (or NaN) seed means the default
                if (seeded) dir = sd;
            }
        }
        if (run_gjk) {
            GjkOut r;
            for (int attempt = 0;; ++attempt) {
                sx_reset(s, cso_from_shapes(seed_pos, ga, gb, dir));
                r = gjk_closest_points(gjk_pos, ga, gb, max_dist, s);
                gi += r.iters;
                // a seeded run that reaches the cap falls back to the unseeded GJK,
                // except for a manifold's cached normal, which parry takes as is
                if (!(seeded && !pfm && attempt == 0 && r.iters == kGjkMaxIters)) break;
                dir = dflt;
            }
            if (is_ball_mode(ball_mode)) {
                if (r.kind == GJK_CLOSEST) {
                    finish_ball_contact(pos1, pos2, pb, radius, prediction, PointProj{false, r.p1}, false,
                                        f3{0.0f, 0.0f, 0.0f}, ball_mode, c);
                } else if (r.kind == GJK_INTERSECTION) {
                    need_epa = true;
                } else {
                    *c = no_contact(kUnsupportedPair);  // parry: unreachable!() in gjk::project_origin
                }
            } else if (r.kind == GJK_CLOSEST) {
                finish_contact(pos1, pos2, pos12, r.p1, r.p2, r.n, is_local_mode(ball_mode), c);
            } else if (r.kind == GJK_NO_INTERSECTION) {
                *c = no_contact(0);
                if (pfm) {  // the manifold caches GJK's separating direction (contact_manifolds_pfm_pfm.rs:148-151)
                    c->n1[0] = r.n.x;
                    c->n1[1] = r.n.y;
                    c->n1[2] = r.n.z;
                }
            } else {
                need_epa = true;
            }
        }
        if (iters) {
            iters[2 * k] = gi;
            iters[2 * k + 1] = 0;
        }
    }
    // Wave-aggregated append to the EPA queue.
    const uint64_t mask = __ballot(need_epa);
    if (mask) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[0], (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
        if (need_epa) {
            EpaItem &it = queue[base + lanes_below(mask)];
            it.pair = pair_id;
            it.dim = s.dim;
            it.mode = ball_mode;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const Cso &v = s.v[q];
                it.v[q][0] = v.p.x; it.v[q][1] = v.p.y; it.v[q][2] = v.p.z;
                it.v[q][3] = v.o1.x; it.v[q][4] = v.o1.y; it.v[q][5] = v.o1.z;
                it.v[q][6] = v.o2.x; it.v[q][7] = v.o2.y; it.v[q][8] = v.o2.z;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 2/3: EPA, G lanes per pair
// ---------------------------------------------------------------------------
// The polytope lives in LDS in compact form: CSO points, faces (vertex and
// neighbour indices, deleted flag, unit normal), the face heap and the
// silhouette stacks.  The shape-space points orig1/orig2 of each CSO vertex
// are only needed for the final witness points, so they go to a per-group
// global buffer; a face's barycentric coordinates are recomputed (the same
// projection as Face::new) when its witness points are built.
// ConvexPolyhedra with at most this many vertices are staged in LDS for the
// EPA support queries; larger ones are read from global memory.
constexpr int kStagedVerts = 64;

template <class I> struct EpaFaceT {
    I pts[3];
    I adj[3];
    uint8_t deleted;
    float n[3];
};

template <int V, int F, int H, int S, int K, class I, int SV> struct EpaState {
    static constexpr int kV = V, kF = F, kH = H, kS = S, kK = K, kSV = SV;
    using Idx = I;
    using Face = EpaFaceT<I>;
    float vp[V][3];
    Face face[F];
    I heap_id[H];
    float heap_nd[H];
    uint32_t sil[S];  // face | opp << 16
    uint32_t stk[K];  // face | opp << 16
    I dl[S];          // faces the current iteration deleted (undone if it cannot commit)
    uint32_t xfer;    // a value one lane passes to its group
    uint8_t vis[F];   // per face: can_be_seen_by the newest vertex, one bit per opposite vertex
    float sv[2][SV > 0 ? SV : 1][3];  // both ConvexPolyhedra's vertices (local frames), if SV > 0
};
// LDS variant (~4.8 KB per pair) sized for the common case; a pair that
// outgrows it re-runs with the large buffers in global memory, which have room
// for parry's 100-iteration cap.
using EpaSmall = EpaState<32, 96, 96, 32, 96, uint8_t, kStagedVerts>;
// The same without staged support vertices, and a smaller one: less LDS per pair,
// more pairs resident per CU (selected by pq_contact_tuning.epa_polytope).
using EpaSmallUnstaged = EpaState<32, 96, 96, 32, 96, uint8_t, 0>;
using EpaTiny = EpaState<24, 64, 64, 24, 64, uint8_t, 0>;
// More faces and heap entries for the same vertices, and a larger polytope
// still on byte ids: fewer pairs move on to the one-wave-per-pair tiers.
using EpaSmallWide = EpaState<32, 160, 160, 32, 96, uint8_t, 0>;
using EpaSmallLarge = EpaState<48, 224, 224, 48, 128, uint8_t, 0>;
// One wave per pair, room for parry's 100 iterations (the vertex count) and for
// the faces a long run creates: ~31 KB and ~81 KB of LDS.
using EpaMedium = EpaState<112, 768, 768, 128, 512, uint16_t, kStagedVerts>;
using EpaBig = EpaState<112, 2048, 2048, 128, 2048, uint16_t, kStagedVerts>;
using EpaLarge = EpaState<112, 8192, 8192, 2048, 8192, uint16_t, 0>;
static_assert(sizeof(EpaFaceT<uint8_t>) == 20, "compact LDS face record");
constexpr uint32_t kOrigFloats = 6;  // orig1, orig2 per CSO vertex

// The argmax reduction over a group of G consecutive lanes: largest value, then
// lowest index (a total order, so any pairing of the lanes gives the same
// winner).  Groups of up to 16 lanes exchange through DPP row permutations --
// mirror within 16 lanes, mirror within 8, then the two quad swaps -- which
// reach every lane of the group in log2(G) steps without the LDS crossbar;
// wider groups use shuffles.  A group's lanes are always active together.
template <int CTRL> __device__ __forceinline__ uint32_t dpp_get(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ void argmax_dpp(float &best, uint32_t &bi, f3 &bp) {
    const float ob = __uint_as_float(dpp_get<CTRL>(__float_as_uint(best)));
    const uint32_t oi = dpp_get<CTRL>(bi);
    const float ox = __uint_as_float(dpp_get<CTRL>(__float_as_uint(bp.x)));
    const float oy = __uint_as_float(dpp_get<CTRL>(__float_as_uint(bp.y)));
    const float oz = __uint_as_float(dpp_get<CTRL>(__float_as_uint(bp.z)));
    if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
        bp = f3{ox, oy, oz};
    }
}
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppQuadXor2 = 0x4E, kDppQuadXor1 = 0xB1;
template <int G> __device__ __forceinline__ void group_argmax(float &best, uint32_t &bi, f3 &bp) {
    if constexpr (G <= 16) {
        if constexpr (G == 16) argmax_dpp<kDppRowMirror>(best, bi, bp);
        if constexpr (G >= 8) argmax_dpp<kDppRowHalfMirror>(best, bi, bp);
        if constexpr (G >= 4) {
            argmax_dpp<kDppQuadXor2>(best, bi, bp);
            argmax_dpp<kDppQuadXor1>(best, bi, bp);
        }
    } else {
        for (int off = G / 2; off >= 1; off >>= 1) {
            const float ob = __shfl_xor(best, off, G);
            const uint32_t oi = __shfl_xor(bi, off, G);
            const float ox = __shfl_xor(bp.x, off, G), oy = __shfl_xor(bp.y, off, G), oz = __shfl_xor(bp.z, off, G);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
                bp = f3{ox, oy, oz};
            }
        }
    }
}

// ConvexPolyhedron support point computed by a group of G consecutive lanes:
// parry takes the first strict maximum of p.d over the vertex list; if vertex
// 0's product is NaN nothing ever beats it, otherwise the result is the lowest
// index among the largest non-NaN products, which is what the group reduction
// (largest value, then lowest index) returns.
template <int G> __device__ __forceinline__ f3 group_support(const ShapeRef &s, f3 d) {
    if (s.type != PQ_SHAPE_CONVEX_POLYHEDRON) return local_support(s, d);
    const float4 p0 = s.v[0];
    const float d0 = dot3(f3{p0.x, p0.y, p0.z}, d);
    if (d0 != d0) return f3{p0.x, p0.y, p0.z};
    const uint32_t sub = __lane_id() % G;
    float best = -INFINITY;
    uint32_t bi = 0xffffffffu;
    f3 bp = f3{p0.x, p0.y, p0.z};  // the coordinates of vertex bi travel with it
    // Each lane scans vertices sub, sub + G, ... in order, four loads in flight.
    for (uint32_t k = sub; k < s.n; k += 4 * G) {
        float4 p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t kk = k + (uint32_t)j * G;
            p[j] = kk < s.n ? s.v[kk] : p0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t kk = k + (uint32_t)j * G;
            const float dd = dot3(f3{p[j].x, p[j].y, p[j].z}, d);
            if (kk < s.n && dd == dd && (bi == 0xffffffffu || dd > best)) {
                best = dd;
                bi = kk;
                bp = f3{p[j].x, p[j].y, p[j].z};
            }
        }
    }
    // Lane sub = 0 always holds a candidate (vertex 0's product is not NaN), so
    // the group's reduction ends on a vertex some lane loaded: no reload.
    group_argmax<G>(best, bi, bp);
    return bp;
}

// The same reduction over vertices staged in LDS (sv = nullptr: not staged).
template <int G> __device__ __forceinline__ f3 group_support_staged(const ShapeRef &s, const float (*sv)[3], f3 d) {
    if (s.type != PQ_SHAPE_CONVEX_POLYHEDRON || sv == nullptr) return group_support<G>(s, d);
    const f3 p0 = f3{sv[0][0], sv[0][1], sv[0][2]};
    const float d0 = dot3(p0, d);
    if (d0 != d0) return p0;
    const uint32_t sub = __lane_id() % G;
    float best = -INFINITY;
    uint32_t bi = 0xffffffffu;
    f3 bp = p0;
    for (uint32_t k = sub; k < s.n; k += G) {
        const f3 q = f3{sv[k][0], sv[k][1], sv[k][2]};
        const float dd = dot3(q, d);
        if (dd == dd && (bi == 0xffffffffu || dd > best)) {
            best = dd;
            bi = k;
            bp = q;
        }
    }
    group_argmax<G>(best, bi, bp);
    return bp;
}

struct Staged {
    const float (*s1)[3];
    const float (*s2)[3];
};

template <int G>
__device__ __forceinline__ Cso group_cso(const iso &pos12, const ShapeRef &g1, const ShapeRef &g2, const Staged &st,
                                         f3 dir) {
    const f3 a = group_support_staged<G>(g1, st.s1, dir);
    const f3 b = g2.type == kShapeOrigin
                     ? pos12.t
                     : iso_point(pos12, group_support_staged<G>(g2, st.s2, qrot_inv(pos12.q, neg3(dir))));
    return Cso{sub3(a, b), a, b};
}

// Copies a ConvexPolyhedron's vertices into LDS (group-cooperative); returns
// the staged array or nullptr if the shape is not staged.
template <int G, class W> __device__ const float (*stage_shape(W *w, int slot, const ShapeRef &s))[3] {
    if (W::kSV == 0 || s.type != PQ_SHAPE_CONVEX_POLYHEDRON || s.n > (uint32_t)W::kSV) return nullptr;
    const uint32_t sub = __lane_id() % G;
    for (uint32_t k = sub; k < s.n; k += G) {
        const float4 p = s.v[k];
        w->sv[slot][k][0] = p.x;
        w->sv[slot][k][1] = p.y;
        w->sv[slot][k][2] = p.z;
    }
    return w->sv[slot];
}

template <class W> __device__ __forceinline__ f3 vpt(const W *w, uint32_t i) {
    return f3{w->vp[i][0], w->vp[i][1], w->vp[i][2]};
}
template <class W> __device__ __forceinline__ void vset(W *w, float *orig, uint32_t i, const Cso &c) {
    w->vp[i][0] = c.p.x;
    w->vp[i][1] = c.p.y;
    w->vp[i][2] = c.p.z;
    float *o = orig + kOrigFloats * i;
    o[0] = c.o1.x; o[1] = c.o1.y; o[2] = c.o1.z;
    o[3] = c.o2.x; o[4] = c.o2.y; o[5] = c.o2.z;
}

// BinaryHeap<FaceId> (max-heap on neg_dist; equal keys compare Equal), with
// std's sift_up / sift_down_to_bottom so equal keys pop in parry's order.
template <class W> struct Heap {
    W *w;
    uint32_t len;
    __device__ int cmp(float a, float b) const { return a < b ? -1 : (a > b ? 1 : 0); }
    __device__ uint32_t sift_up(uint32_t start, uint32_t pos) {
        const uint32_t eid = w->heap_id[pos];
        const float end = w->heap_nd[pos];
        while (pos > start) {
            const uint32_t parent = (pos - 1) / 2;
            if (cmp(end, w->heap_nd[parent]) <= 0) break;
            w->heap_id[pos] = w->heap_id[parent];
            w->heap_nd[pos] = w->heap_nd[parent];
            pos = parent;
        }
        w->heap_id[pos] = (typename W::Idx)eid;
        w->heap_nd[pos] = end;
        return pos;
    }
    // Places an element at position pos (the heap's current end plus the pushes
    // before it) and sifts it up; the caller advances len.
    __device__ void insert_at(uint32_t pos, uint32_t id, float nd) {
        w->heap_id[pos] = (typename W::Idx)id;
        w->heap_nd[pos] = nd;
        sift_up(0, pos);
    }
    __device__ bool push(uint32_t id, float nd) {
        if (len >= (uint32_t)W::kH) return false;
        w->heap_id[len] = (typename W::Idx)id;
        w->heap_nd[len] = nd;
        ++len;
        sift_up(0, len - 1);
        return true;
    }
    __device__ void sift_down_to_bottom(uint32_t pos) {
        const uint32_t end = len, start = pos;
        const uint32_t eid = w->heap_id[pos];
        const float end_nd = w->heap_nd[pos];
        uint32_t child = 2 * pos + 1;
        while (child + 2 <= end) {
            if (cmp(w->heap_nd[child], w->heap_nd[child + 1]) <= 0) child += 1;
            w->heap_id[pos] = w->heap_id[child];
            w->heap_nd[pos] = w->heap_nd[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (child == end - 1) {
            w->heap_id[pos] = w->heap_id[child];
            w->heap_nd[pos] = w->heap_nd[child];
            pos = child;
        }
        w->heap_id[pos] = (typename W::Idx)eid;
        w->heap_nd[pos] = end_nd;
        sift_up(start, pos);
    }
    __device__ bool pop(uint32_t &id, float &nd) {
        if (len == 0) return false;
        --len;
        uint32_t iid = w->heap_id[len];
        float ind = w->heap_nd[len];
        if (len > 0) {
            const uint32_t t_id = w->heap_id[0];
            const float t_nd = w->heap_nd[0];
            w->heap_id[0] = (typename W::Idx)iid;
            w->heap_nd[0] = ind;
            iid = t_id;
            ind = t_nd;
            sift_down_to_bottom(0);
        }
        id = iid;
        nd = ind;
        return true;
    }
};

enum : uint32_t { EPA_OK = 0, EPA_NONE = 1, EPA_OVERFLOW = 2, EPA_RUNNING = 3 };

// Barycentric coordinates of Face::new's projection (epa3.rs:96-118,
// TrianglePointLocation::barycentric_coordinates); returns proj_inside.
__device__ __forceinline__ bool face_bcoords(f3 pa, f3 pb, f3 pc, float bc[3]) {
    Loc l;
    bool inside;
    const f3 proj = tri_origin(pa, pb, pc, l, inside);
    bc[0] = bc[1] = bc[2] = 0.0f;
    const float e = kEps * 100.0f;
    if (l.kind == LOC_VERTEX) {
        bc[0] = l.id == 0 ? 1.0f : 0.0f;
        bc[1] = l.id == 1 ? 1.0f : 0.0f;
        bc[2] = l.id == 2 ? 1.0f : 0.0f;
        return inside || norm_sq3(proj) < e * e;
    }
    if (l.kind == LOC_EDGE) {  // edges (0,1), (1,2), (0,2)
        bc[0] = l.id == 1 ? 0.0f : l.b0;
        bc[1] = l.id == 0 ? l.b1 : (l.id == 1 ? l.b0 : 0.0f);
        bc[2] = l.id == 0 ? 0.0f : l.b1;
        return inside || norm_sq3(proj) < e * e;
    }
    if (l.kind == LOC_FACE) {
        bc[0] = l.b0;
        bc[1] = l.b1;
        bc[2] = l.b2;
        return true;
    }
    return false;
}

// ccw_face_normal (utils/ccw_face_normal.rs:19-27), zero if degenerate.
__device__ __forceinline__ f3 ccw_normal(f3 pa, f3 pb, f3 pc) {
    const f3 r = cross3(sub3(pb, pa), sub3(pc, pa));
    const float sq = norm_sq3(r);
    f3 n = f3{0.0f, 0.0f, 0.0f};
    if (sq > kEps * kEps) n = div3(r, sqrtf(sq));
    return n;
}

// Face::new: fills face f; returns proj_inside.
template <class W>
__device__ bool face_new(W *w, uint32_t f, uint32_t a, uint32_t b, uint32_t c, uint32_t j0, uint32_t j1, uint32_t j2) {
    using I = typename W::Idx;
    typename W::Face &F = w->face[f];
    F.pts[0] = (I)a; F.pts[1] = (I)b; F.pts[2] = (I)c;
    F.adj[0] = (I)j0; F.adj[1] = (I)j1; F.adj[2] = (I)j2;
    F.deleted = 0;
    const f3 pa = vpt(w, a), pb = vpt(w, b), pc = vpt(w, c);
    float bc[3];
    const bool in = face_bcoords(pa, pb, pc, bc);
    const f3 n = ccw_normal(pa, pb, pc);
    F.n[0] = n.x; F.n[1] = n.y; F.n[2] = n.z;
    return in;
}

template <class W> __device__ __forceinline__ f3 fnormal(const W *w, uint32_t f) {
    return f3{w->face[f].n[0], w->face[f].n[1], w->face[f].n[2]};
}

// Face::closest_points (epa3.rs:120-130).
template <class W> __device__ void face_points(const W *w, const float *orig, uint32_t f, f3 &p1, f3 &p2) {
    const typename W::Face &F = w->face[f];
    float bc[3];
    face_bcoords(vpt(w, F.pts[0]), vpt(w, F.pts[1]), vpt(w, F.pts[2]), bc);
    f3 o1[3], o2[3];
    for (int q = 0; q < 3; ++q) {
        const float *o = orig + kOrigFloats * F.pts[q];
        o1[q] = f3{o[0], o[1], o[2]};
        o2[q] = f3{o[3], o[4], o[5]};
    }
    p1 = add3(add3(mul3(o1[0], bc[0]), mul3(o1[1], bc[1])), mul3(o1[2], bc[2]));
    p2 = add3(add3(mul3(o2[0], bc[0]), mul3(o2[1], bc[1])), mul3(o2[2], bc[2]));
}

// Element k (0, 1 or 2) of three values, by selects rather than an indexed array.
__device__ __forceinline__ uint32_t pick3(uint32_t k, uint32_t a, uint32_t b, uint32_t c) {
    return k == 0 ? a : (k == 1 ? b : c);
}

// Face::next_ccw_pt_id (epa3.rs:135-151) on a face record held in registers.
template <class Fc> __device__ __forceinline__ uint32_t ccw_index(const Fc &f, uint32_t id) {
    if (f.pts[0] == id) return 1;
    if (f.pts[1] == id) return 2;
    return 0;
}

template <class W> __device__ __forceinline__ uint32_t next_ccw(const W *w, uint32_t f, uint32_t id) {
    if (w->face[f].pts[0] == id) return 1;
    if (w->face[f].pts[1] == id) return 2;
    return 0;
}

// Face::can_be_seen_by (epa3.rs:147-163).
template <class W> __device__ bool seen_by(const W *w, uint32_t f, uint32_t point, uint32_t opp) {
    const typename W::Face &F = w->face[f];
    const f3 p0 = vpt(w, F.pts[opp]);
    const f3 p1 = vpt(w, F.pts[(opp + 1) % 3]);
    const f3 p2 = vpt(w, F.pts[(opp + 2) % 3]);
    const f3 pt = vpt(w, point);
    if (dot3(sub3(pt, p0), fnormal(w, f)) >= -kEpsTol) return true;
    const float E = kEps * 100.0f;
    return norm_sq3(cross3(sub3(p2, p1), sub3(pt, p1))) <= E * E;  // Triangle::is_affinely_dependent
}

// A first-tier EPA polytope at the start of an iteration, handed to the
// wave-per-pair tier so that the pair continues where it stopped instead of
// starting over (byte ids: at most 48 vertices, 224 faces, 224 heap entries).
constexpr int kContV = 48, kContF = 224, kContH = 224;
constexpr uint32_t kNoCont = 0xffffffffu;
struct EpaCont {
    uint32_t nv, nf, len, best_id, niter, pad[3];
    float max_dist, old_dist, padf[2];
    float vp[kContV][3];
    float orig[kContV][6];
    uint32_t face[kContF][2];  // pts | deleted << 24, adj
    uint32_t heap_id[kContH];
    float heap_nd[kContH];
};
// Where a tier writes the polytopes it hands on (cap 0: it hands on none).
struct ContSink {
    EpaCont *recs;
    uint32_t *count;
    uint32_t cap;
};
struct EpaScalars {
    uint32_t nv, nf, best_id, niter;
    float max_dist, old_dist;
};

// Writes the group's polytope to a continuation record; returns its index or
// kNoCont (no room, or a polytope type the record cannot hold).
template <int G, class W>
__device__ uint32_t epa_export(W *w, const float *orig, const ContSink &sink, const EpaScalars &x, uint32_t len) {
    if constexpr (sizeof(typename W::Idx) == 1 && W::kV <= kContV && W::kF <= kContF && W::kH <= kContH) {
        if (sink.cap == 0) return kNoCont;
        const uint32_t sub = __lane_id() % G;
        if (sub == 0) w->xfer = atomicAdd(sink.count, 1u);
        // orig rows were written to global memory by this wave; the group's
        // lanes read each other's rows below
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const uint32_t ci = w->xfer;
        if (ci >= sink.cap) return kNoCont;
        EpaCont *r = sink.recs + ci;
        if (sub == 0) {
            r->nv = x.nv;
            r->nf = x.nf;
            r->len = len;
            r->best_id = x.best_id;
            r->niter = x.niter;
            r->max_dist = x.max_dist;
            r->old_dist = x.old_dist;
        }
        for (uint32_t k = sub; k < x.nv; k += G) {
            for (int q = 0; q < 3; ++q) r->vp[k][q] = w->vp[k][q];
            for (int q = 0; q < 6; ++q) r->orig[k][q] = orig[kOrigFloats * k + q];
        }
        for (uint32_t k = sub; k < x.nf; k += G) {
            const typename W::Face &F = w->face[k];
            r->face[k][0] = (uint32_t)F.pts[0] | ((uint32_t)F.pts[1] << 8) | ((uint32_t)F.pts[2] << 16) |
                            ((uint32_t)F.deleted << 24);
            r->face[k][1] = (uint32_t)F.adj[0] | ((uint32_t)F.adj[1] << 8) | ((uint32_t)F.adj[2] << 16);
        }
        for (uint32_t k = sub; k < len; k += G) {
            r->heap_id[k] = w->heap_id[k];
            r->heap_nd[k] = w->heap_nd[k];
        }
        return ci;
    } else {
        return kNoCont;
    }
}

// Loads a continuation record into the group's polytope; face normals are
// recomputed from the vertices exactly as Face::new computed them.
template <int G, class W>
__device__ void epa_import(W *w, float *orig, const EpaCont *r, EpaScalars &x, uint32_t &len) {
    if constexpr (W::kV >= kContV && W::kF >= kContF && W::kH >= kContH) {
        const uint32_t sub = __lane_id() % G;
        x.nv = r->nv;
        x.nf = r->nf;
        len = r->len;
        x.best_id = r->best_id;
        x.niter = r->niter;
        x.max_dist = r->max_dist;
        x.old_dist = r->old_dist;
        for (uint32_t k = sub; k < x.nv; k += G) {
            for (int q = 0; q < 3; ++q) w->vp[k][q] = r->vp[k][q];
            for (int q = 0; q < 6; ++q) orig[kOrigFloats * k + q] = r->orig[k][q];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        using I = typename W::Idx;
        for (uint32_t k = sub; k < x.nf; k += G) {
            const uint32_t a = r->face[k][0], b = r->face[k][1];
            typename W::Face &F = w->face[k];
            F.pts[0] = (I)(a & 0xffu);
            F.pts[1] = (I)((a >> 8) & 0xffu);
            F.pts[2] = (I)((a >> 16) & 0xffu);
            F.deleted = (uint8_t)(a >> 24);
            F.adj[0] = (I)(b & 0xffu);
            F.adj[1] = (I)((b >> 8) & 0xffu);
            F.adj[2] = (I)((b >> 16) & 0xffu);
            const f3 nn = ccw_normal(vpt(w, F.pts[0]), vpt(w, F.pts[1]), vpt(w, F.pts[2]));
            F.n[0] = nn.x;
            F.n[1] = nn.y;
            F.n[2] = nn.z;
        }
        for (uint32_t k = sub; k < len; k += G) {
            w->heap_id[k] = (I)r->heap_id[k];
            w->heap_nd[k] = r->heap_nd[k];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
}

// EPA::closest_points (epa3.rs:428-651) for one pair, executed by a group of G
// lanes one iteration at a time: epa_begin, then epa_step until it returns a
// result (epa_run below runs the whole loop).
//
// Every iteration is checkpointable: the heap's top is only peeked, the new
// vertex goes to the first free slot without being counted, the best face and
// distances are kept aside, and the faces the iteration deletes are listed.
// Only when the polytope is known to have room for the iteration's new faces
// and heap entries is the iteration committed -- the top popped (nothing reads
// the heap between parry's pop and its pushes, so the heap ends up the same),
// the vertex counted and the deletions kept.  A polytope that runs out of room
// is therefore still the one the iteration started with: the deletions are
// undone and, if `sink` has room, the polytope is written out for the next
// tier to continue (`cont` gets the record's index).  `resume` starts from such
// a record instead of the GJK simplex.
template <class W> struct EpaRegs {
    Heap<W> heap;  // the polytope's face heap (LDS) and its length
    uint32_t nv, nf, best_id, niter;
    float max_dist, old_dist;
};

// The set-up before EPA's first iteration: the polytope from the GJK simplex, or
// from a continuation record.  Returns EPA_RUNNING, or the result if EPA ends
// here.
template <int G, class W>
__device__ uint32_t epa_begin(W *w, float *orig, const iso &pos12, const ShapeRef &g1, const ShapeRef &g2,
                              const EpaItem *it, const EpaCont *resume, Staged &st, EpaRegs<W> &r, f3 &p1, f3 &p2,
                              f3 &n, uint32_t &epa_iters, uint32_t &cont) {
    st = Staged{stage_shape<G>(w, 0, g1), stage_shape<G>(w, 1, g2)};
    // The group's lanes wrote the staged vertices; make them visible to the
    // whole wave before any lane reads another lane's entries.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    r = EpaRegs<W>{Heap<W>{w, 0}, 0, 0, 0, 0, kFltMax, 0.0f};
    Heap<W> &heap = r.heap;
    uint32_t &nv = r.nv, &nf = r.nf, &best_id = r.best_id, &niter = r.niter;
    float &max_dist = r.max_dist, &old_dist = r.old_dist;
    epa_iters = 0;
    cont = kNoCont;
    if (resume) {
        EpaScalars x;
        epa_import<G>(w, orig, resume, x, heap.len);
        nv = x.nv;
        nf = x.nf;
        best_id = x.best_id;
        niter = x.niter;
        max_dist = x.max_dist;
        old_dist = x.old_dist;
    } else {
        const uint32_t dim = it->dim;
        for (uint32_t i = 0; i <= dim; ++i) {
            w->vp[i][0] = it->v[i][0];
            w->vp[i][1] = it->v[i][1];
            w->vp[i][2] = it->v[i][2];
            for (int q = 0; q < 6; ++q) orig[kOrigFloats * i + q] = it->v[i][3 + q];
        }
        nv = dim + 1;
        if (dim == 0) {
            p1 = p2 = f3{0.0f, 0.0f, 0.0f};
            n = f3{0.0f, 1.0f, 0.0f};
            return EPA_OK;
        } else if (dim == 3) {
            const f3 dp1 = sub3(vpt(w, 1), vpt(w, 0));
            const f3 dp2 = sub3(vpt(w, 2), vpt(w, 0));
            const f3 dp3 = sub3(vpt(w, 3), vpt(w, 0));
            if (dot3(cross3(dp1, dp2), dp3) > 0.0f) {
                for (int q = 0; q < 3; ++q) {
                    const float t = w->vp[1][q];
                    w->vp[1][q] = w->vp[2][q];
                    w->vp[2][q] = t;
                }
                for (int q = 0; q < 6; ++q) {
                    const float t = orig[kOrigFloats * 1 + q];
                    orig[kOrigFloats * 1 + q] = orig[kOrigFloats * 2 + q];
                    orig[kOrigFloats * 2 + q] = t;
                }
            }
            const bool in0 = face_new(w, 0, 0, 1, 2, 3, 1, 2);
            const bool in1 = face_new(w, 1, 1, 3, 2, 3, 2, 0);
            const bool in2 = face_new(w, 2, 0, 2, 3, 0, 1, 3);
            const bool in3 = face_new(w, 3, 0, 3, 1, 2, 1, 0);
            nf = 4;
            const bool ins[4] = {in0, in1, in2, in3};
    #pragma unroll
            for (uint32_t f = 0; f < 4; ++f) {
                if (ins[f]) {
                    const float dist = dot3(fnormal(w, f), vpt(w, f));
                    if (-dist > kEpsTol) return EPA_NONE;  // FaceId::new -> None
                    heap.push(f, -dist);
                }
            }
            if (!(in0 || in1 || in2 || in3)) return EPA_NONE;
        } else {
            if (dim == 1) {
                // orthonormal_subspace_basis([dpt]): first vector a x dpt
                const f3 dpt = sub3(vpt(w, 1), vpt(w, 0));
                f3 a = fabsf(dpt.x) > fabsf(dpt.y) ? f3{dpt.z, 0.0f, -dpt.x} : f3{0.0f, -dpt.z, dpt.y};
                a = div3(a, sqrtf(norm_sq3(a)));
                vset(w, orig, 2, group_cso<G>(pos12, g1, g2, st, cross3(a, dpt)));
                nv = 3;
            }
            face_new(w, 0, 0, 1, 2, 1, 1, 1);
            face_new(w, 1, 0, 2, 1, 0, 0, 0);
            nf = 2;
            heap.push(0, 0.0f);
            heap.push(1, 0.0f);
        }
        if (heap.len == 0) return EPA_NONE;
        best_id = w->heap_id[0];
    }
    return EPA_RUNNING;
}

// One EPA iteration (the body of the loop of epa3.rs:499-647).  Returns
// EPA_RUNNING while EPA goes on, else the result.
template <int G, class W>
__device__ uint32_t epa_step(W *w, float *orig, const iso &pos12, const ShapeRef &g1, const ShapeRef &g2,
                             const Staged &st, const ContSink &sink, EpaRegs<W> &r, f3 &p1, f3 &p2, f3 &n,
                             uint32_t &epa_iters, uint32_t &cont) {
    const float eps_tol = kEps * 100.0f;
    Heap<W> &heap = r.heap;
    uint32_t &nv = r.nv, &nf = r.nf, &best_id = r.best_id, &niter = r.niter;
    float &max_dist = r.max_dist, &old_dist = r.old_dist;
    // parry's loop ends when the heap is empty or after 100 iterations
    auto done = [&]() -> uint32_t {
        epa_iters = niter + 1;
        face_points(w, orig, best_id, p1, p2);
        n = fnormal(w, best_id);
        return EPA_OK;
    };
    // Hands the polytope as it was at the start of this iteration to the next tier.
    auto overflow = [&](uint32_t ndel, uint32_t fid) -> uint32_t {
        for (uint32_t k = 0; k < ndel; ++k) w->face[w->dl[k]].deleted = 0;
        w->face[fid].deleted = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        cont = epa_export<G>(w, orig, sink, EpaScalars{nv, nf, best_id, niter, max_dist, old_dist}, heap.len);
        return EPA_OVERFLOW;
    };
    // Faces deleted since they were pushed are popped and skipped here, as
    // parry's loop skips them (`continue`), rather than one per call.
    while (heap.len > 0 && w->face[w->heap_id[0]].deleted) {
        uint32_t pid;
        float pnd;
        heap.pop(pid, pnd);
    }
    if (heap.len == 0) return done();
    {
        const uint32_t fid = w->heap_id[0];
        const float fnd = w->heap_nd[0];
        const typename W::Face face = w->face[fid];
        const f3 fn = f3{face.n[0], face.n[1], face.n[2]};
        const Cso c = group_cso<G>(pos12, g1, g2, st, fn);
        if (nv >= (uint32_t)W::kV) return overflow(0, fid);
        const uint32_t sp = nv;  // counted when the iteration commits
        vset(w, orig, sp, c);
        const float cand = dot3(c.p, fn);
        uint32_t best_t = best_id;
        float max_t = max_dist;
        if (cand < max_t) {
            best_t = fid;
            max_t = cand;
        }
        const float curr_dist = -fnd;
        if (max_t - curr_dist < eps_tol || (fabsf(curr_dist - old_dist) < kEps && cand < max_t)) {
            face_points(w, orig, best_t, p1, p2);
            n = fnormal(w, best_t);
            epa_iters = niter + 1;
            return EPA_OK;
        }
        w->face[fid].deleted = 1;
        // Face::can_be_seen_by (epa3.rs:153-166) of every live face from the new
        // vertex, for each of the three opposite vertices, by the group's lanes in
        // parallel; the depth-first walk below only reads the bits.
        {
            const uint32_t sub = __lane_id() % G;
            const f3 pt = vpt(w, sp);
            const float E = kEps * 100.0f;
            // Deleted faces get bits too (the walk never reads them): with no
            // branch in the body, the loads of consecutive faces overlap.
#pragma unroll 2
            for (uint32_t f = sub; f < nf; f += G) {
                const typename W::Face F = w->face[f];
                const f3 q0 = vpt(w, F.pts[0]), q1 = vpt(w, F.pts[1]), q2 = vpt(w, F.pts[2]);
                const f3 fnn = f3{F.n[0], F.n[1], F.n[2]};
                const f3 qs[3] = {q0, q1, q2};
                uint32_t bits = 0;
#pragma unroll
                for (int opp = 0; opp < 3; ++opp) {
                    const f3 p0 = qs[opp], p1 = qs[(opp + 1) % 3], p2 = qs[(opp + 2) % 3];
                    const bool front = dot3(sub3(pt, p0), fnn) >= -kEpsTol;
                    const bool flat = norm_sq3(cross3(sub3(p2, p1), sub3(pt, p1))) <= E * E;  // is_affinely_dependent
                    bits |= (uint32_t)(front | flat) << opp;
                }
                w->vis[f] = (uint8_t)bits;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        // compute_silhouette (epa3.rs:653-675) for the three neighbours in order
        // adj[0], adj[1], adj[2]; a visible face recurses into adj[(opp + 2) % 3],
        // then adj[opp].  The recursion is an explicit stack of face | opp << 16
        // whose top entry lives in a register.  After a visible face, the face
        // visited next (its first far neighbour) is loaded together with the
        // second one's points and carried in registers, so a step costs one
        // round trip to LDS.  Deleted faces are listed in dl for an undo.
        uint32_t ns = 0, sk = 0, depth = 0, top = 0, ndel = 0;
        typename W::Face cf = w->face[face.adj[0]];
        uint32_t cvis = w->vis[face.adj[0]];
        uint32_t cur = face.adj[0] | (ccw_index(cf, face.pts[0]) << 16);
        {
            const uint32_t o2 = next_ccw(w, face.adj[2], face.pts[2]);
            const uint32_t o1 = next_ccw(w, face.adj[1], face.pts[1]);
            w->stk[0] = face.adj[2] | (o2 << 16);
            top = face.adj[1] | (o1 << 16);
            sk = 1;
            depth = 2;
        }
        // One latch: both ways on (into a deleted face's first far neighbour, or
        // back to the stack's top) end in the same read of the next face.  With
        // the stack path as a separate branch the compiler split the walk into
        // nested loops, and a wave's groups that had to pop waited for the others.
        for (;;) {
            const uint32_t id = cur & 0xffffu, opp = cur >> 16;
            const bool del = !cf.deleted && ((cvis >> opp) & 1u);
            if (!cf.deleted && !del) {
                if (ns >= (uint32_t)W::kS) return overflow(ndel, fid);
                w->sil[ns++] = cur;
            }
            if (!del && depth == 0) break;
            uint32_t nid, pt1 = 0, popped = 0;
            if (del) {
                if (ndel >= (uint32_t)W::kS) return overflow(ndel, fid);
                w->dl[ndel++] = (typename W::Idx)id;
                w->face[id].deleted = 1;
                // a1 = (opp + 2) % 3 and a2 = opp, selected by value so the
                // record stays in registers
                const uint32_t adj1 = pick3(opp, cf.adj[2], cf.adj[0], cf.adj[1]);
                const uint32_t adj2 = pick3(opp, cf.adj[0], cf.adj[1], cf.adj[2]);
                pt1 = pick3(opp, cf.pts[2], cf.pts[0], cf.pts[1]);
                const uint32_t pt2 = pick3(opp, cf.pts[0], cf.pts[1], cf.pts[2]);
                const uint32_t o2 = next_ccw(w, adj2, pt2);
                if (depth > 0) {
                    if (sk >= (uint32_t)W::kK) return overflow(ndel, fid);
                    w->stk[sk++] = top;
                }
                top = adj2 | (o2 << 16);
                ++depth;
                nid = adj1;
            } else {
                popped = top;
                nid = top & 0xffffu;
                --depth;
                const uint32_t below = w->stk[sk != 0u ? sk - 1u : 0u];  // slot 0 always exists
                if (depth > 0) {
                    top = below;
                    --sk;
                }
            }
            const typename W::Face f = w->face[nid];
            cvis = w->vis[nid];
            cur = del ? (nid | (ccw_index(f, pt1) << 16)) : popped;
            cf = f;
            if (del && nid == id) cf.deleted = 1;  // a face adjacent to itself (degenerate ring)
        }
        const uint32_t first_new = nf;
        if (ns == 0) return EPA_NONE;
        const uint32_t sub = __lane_id() % G;
        const uint32_t gbase = __lane_id() - sub;
        const uint64_t gmask = G >= 64 ? ~0ull : ((1ull << (G & 63)) - 1ull);
        const uint64_t below = (1ull << sub) - 1ull;
        // Room for the new faces and their heap entries (at most one per live
        // silhouette edge), checked before anything is committed.
        uint32_t n_live = 0;
        for (uint32_t c0 = 0; c0 < ns; c0 += G) {
            const uint32_t q = c0 + sub;
            const bool alive = q < ns && !w->face[w->sil[q] & 0xffffu].deleted;
            n_live += (uint32_t)__popcll((__ballot(alive) >> gbase) & gmask);
        }
        if (first_new + n_live > (uint32_t)W::kF || heap.len - 1 + n_live > (uint32_t)W::kH)
            return overflow(ndel, fid);
        // Commit: parry popped this face before all of the above.
        {
            uint32_t pid;
            float pnd;
            heap.pop(pid, pnd);
        }
        ++nv;
        best_id = best_t;
        max_dist = max_t;
        old_dist = curr_dist;
        // The new faces (epa3.rs:598-626), G silhouette edges at a time, one per
        // lane.  The edges whose face is still live get consecutive ids in
        // silhouette order (a ballot gives each its rank), and each lane builds
        // its edge's face (Face::new reads vertex positions only, and each edge
        // rewrites its own adjacency slot).  Parry's loop then either returns at
        // the first new face that ends EPA -- its heap pushes before that would
        // be discarded anyway -- or pushes every new face in silhouette order;
        // the lane holding an edge's face does that edge's push.
        uint32_t n_alive = 0;
        for (uint32_t c0 = 0; c0 < ns; c0 += G) {
            const uint32_t q = c0 + sub;
            uint32_t e = 0;
            bool alive = false;
            if (q < ns) {
                e = w->sil[q];
                alive = !w->face[e & 0xffffu].deleted;
            }
            const uint64_t am = (__ballot(alive) >> gbase) & gmask;
            const uint32_t new_id = first_new + n_alive + (uint32_t)__popcll(am & below);
            bool in = false;
            float dist = 0.0f;
            if (alive) {
                const uint32_t ef = e & 0xffffu, opp = e >> 16;
                const uint32_t pt1 = w->face[ef].pts[(opp + 2) % 3];
                const uint32_t pt2 = w->face[ef].pts[(opp + 1) % 3];
                in = face_new(w, new_id, pt1, pt2, sp, ef, new_id + 1, new_id - 1);
                w->face[ef].adj[(opp + 1) % 3] = (typename W::Idx)new_id;
                if (in) dist = dot3(fnormal(w, new_id), vpt(w, pt1));
            }
            n_alive += (uint32_t)__popcll(am);
            const bool pushed = alive && in;
            const bool closer = pushed && dist < curr_dist;
            const bool stop = closer || (pushed && -dist > kEpsTol);
            const uint64_t stop_m = (__ballot(stop) >> gbase) & gmask;
            const uint64_t closer_m = (__ballot(closer) >> gbase) & gmask;
            if (stop_m != 0) {
                if ((closer_m >> __builtin_ctzll(stop_m)) & 1ull) {
                    face_points(w, orig, fid, p1, p2);  // the popped face: its points are unchanged
                    n = fn;
                    epa_iters = niter + 1;
                    return EPA_OK;
                }
                return EPA_NONE;  // FaceId::new -> None
            }
            const uint64_t push_m = (__ballot(pushed) >> gbase) & gmask;
            for (uint64_t m = push_m; m != 0; m &= m - 1) {
                if (sub == (uint32_t)__builtin_ctzll(m))
                    heap.insert_at(heap.len + (uint32_t)__popcll(push_m & below), new_id, -dist);
            }
            heap.len += (uint32_t)__popcll(push_m);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        nf = first_new + n_alive;
        if (first_new == nf) return EPA_NONE;
        w->face[first_new].adj[2] = (typename W::Idx)(nf - 1);
        w->face[nf - 1].adj[1] = (typename W::Idx)first_new;
        ++niter;
        if (niter > 100) return done();
    }
    return EPA_RUNNING;
}

// EPA::closest_points (epa3.rs:428-651) for one pair by a group of G lanes.
template <int G, class W>
__device__ uint32_t epa_run(W *w, float *orig, const iso &pos12, const ShapeRef &g1, const ShapeRef &g2,
                            const EpaItem *it, const EpaCont *resume, const ContSink &sink, f3 &p1, f3 &p2, f3 &n,
                            uint32_t &epa_iters, uint32_t &cont) {
    Staged st;
    EpaRegs<W> r;
    uint32_t s = epa_begin<G>(w, orig, pos12, g1, g2, it, resume, st, r, p1, p2, n, epa_iters, cont);
    while (s == EPA_RUNNING) s = epa_step<G>(w, orig, pos12, g1, g2, st, sink, r, p1, p2, n, epa_iters, cont);
    return s;
}


// The EPA problem of a queued pair (contact_support_map_support_map.rs:60-76):
// its two shapes (their feature maps in kPairFeatureMaps mode) and relative pose,
// or, for a Ball, the other shape against the ball centre
// (point_support_map.rs:39-49).
struct EpaProblem {
    iso pos;
    ShapeRef a, b;
};

__device__ EpaProblem epa_problem(const ShapeRecord *rec, const float4 *verts, const float *poses,
                                  const uint32_t *pairs, const EpaItem *it) {
    const uint32_t k = it->pair;
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
    const iso pos12 = iso_inv_mul(pos1, pos2);
    ShapeRef g1 = load_shape(rec, verts, i), g2 = load_shape(rec, verts, j);
    const uint32_t mode = it->mode;
    if (mode == kPairFeatureMaps) {
        float br;
        (void)feature_map(g1, br);
        (void)feature_map(g2, br);
    }
    if (is_ball_mode(mode)) {
        if (mode == kBallFirst) return EpaProblem{point_proj_iso(iso_inverse(pos12).t), g2, origin_shape()};
        return EpaProblem{point_proj_iso(pos12.t), g1, origin_shape()};
    }
    return EpaProblem{pos12, g1, g2};
}

// Writes the result of a queued pair's EPA (one lane per pair): the contact, or
// the pair's entry in the next tier's list if its polytope outgrew this tier's.
__device__ void epa_finish(const ShapeRecord *rec, const float4 *verts, const float *poses, const uint32_t *pairs,
                           float prediction, const EpaItem *it, uint32_t st, f3 p1, f3 p2, f3 n, uint32_t ei,
                           uint32_t cont, DevContact *out, uint32_t *iters, uint32_t *overflow_list,
                           uint32_t *overflow_cont, uint32_t *overflow_count, uint32_t slot) {
    const uint32_t k = it->pair;
    if (st == EPA_OVERFLOW) {
        if (overflow_list) {
            const uint32_t pos = atomicAdd(overflow_count, 1u);
            overflow_list[pos] = slot;
            if (overflow_cont) overflow_cont[pos] = cont;
        } else {
            out[k] = no_contact(kUnsupportedPair);  // cannot happen with the large buffers
        }
        return;
    }
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
    const iso pos12 = iso_inv_mul(pos1, pos2);
    const uint32_t mode = it->mode;
    if (is_ball_mode(mode)) {
        // the Ball-vs-convex contact from the projection of the ball centre
        const iso pb = mode == kBallFirst ? iso_inverse(pos12) : pos12;
        const float radius = rec[mode == kBallFirst ? i : j].param[0];
        const PointProj proj{true, st == EPA_OK ? p1 : pb.t};
        finish_ball_contact(pos1, pos2, pb, radius, prediction, proj, false, f3{0.0f, 0.0f, 0.0f}, mode, out + k);
    } else if (st == EPA_OK) {
        finish_contact(pos1, pos2, pos12, p1, p2, n, is_local_mode(mode), out + k);
    } else {
        out[k] = no_contact(0);  // EPA failed: NoIntersection (contact_support_map_support_map.rs:75-76)
        if (mode == kPairFeatureMaps) out[k].n1[0] = 1.0f;  // ... of Vector::x_axis()
    }
    if (iters) iters[2 * k + 1] = ei;
}

template <int G, class W>
__device__ void epa_item(W *w, float *orig, const ShapeRecord *rec, const float4 *verts, const float *poses,
                         const uint32_t *pairs, float prediction, const EpaItem *it, const EpaCont *resume,
                         const ContSink &sink, DevContact *out, uint32_t *iters, uint32_t *overflow_list,
                         uint32_t *overflow_cont, uint32_t *overflow_count, uint32_t slot) {
    const EpaProblem pr = epa_problem(rec, verts, poses, pairs, it);
    f3 p1, p2, n;
    uint32_t ei = 0, cont = kNoCont;
    const uint32_t st = epa_run<G>(w, orig, pr.pos, pr.a, pr.b, it, resume, sink, p1, p2, n, ei, cont);
    if (__lane_id() % G != 0) return;
    epa_finish(rec, verts, poses, pairs, prediction, it, st, p1, p2, n, ei, cont, out, iters, overflow_list,
               overflow_cont, overflow_count, slot);
}

// EPA tiers.  Each reads a list of queued pairs (the queue itself for the first
// tier) and sends the pairs whose polytope outgrows its buffers to the next
// tier's list: G lanes per pair with a small LDS polytope, then one wave per
// pair with medium and big LDS polytopes, then a global-memory polytope.
// At most 168 VGPRs (three waves per SIMD) where a wave's polytopes take at
// most 16 KB of LDS, which leaves room for that many waves: the first tier's
// latency-bound code needs them.  Larger polytopes are limited by LDS anyway,
// and a wave per pair (G = 64) keeps its registers.
template <int G, class W> constexpr int epa_min_waves() { return G < 64 && (64 / G) * sizeof(W) <= 16384 ? 3 : 1; }
template <int G, class W>
__global__ __launch_bounds__(64, (epa_min_waves<G, W>())) void epa_tier_kernel(const ShapeRecord *__restrict__ rec,
                                                      const float4 *__restrict__ verts, const float *__restrict__ poses,
                                                      const uint32_t *__restrict__ pairs, float prediction,
                                                      const EpaItem *__restrict__ queue, const uint32_t *in_list,
                                                      const uint32_t *in_count, uint32_t *next,
                                                      const uint32_t *in_cont, const EpaCont *cont_recs,
                                                      ContSink sink, DevContact *__restrict__ out,
                                                      uint32_t *__restrict__ iters, uint32_t *out_list,
                                                      uint32_t *out_cont, uint32_t *out_count, float *orig_scratch) {
    constexpr int kGroups = 64 / G;
    __shared__ W w[kGroups];
    const uint32_t group = threadIdx.x / G;
    const uint32_t gid = blockIdx.x * kGroups + group;
    float *orig = orig_scratch + (size_t)gid * W::kV * kOrigFloats;
    const uint32_t count = __hip_atomic_load(in_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // The wave takes its next kGroups list entries from the tier's counter
    // whenever all its groups are done, so no wave is left with a long run of
    // long pairs.  The fetch happens where the whole wave is converged and the
    // exit test is wave-uniform: no cross-lane exchange under divergence.
    for (;;) {
        uint32_t base = 0;
        if (__lane_id() == 0) base = atomicAdd(next, (uint32_t)kGroups);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= count) break;
        const uint32_t q = base + group;
        if (q < count) {
            const uint32_t slot = in_list ? in_list[q] : q;
            const uint32_t ci = in_cont ? in_cont[q] : kNoCont;
            epa_item<G>(&w[group], orig, rec, verts, poses, pairs, prediction, queue + slot,
                        ci != kNoCont ? cont_recs + ci : nullptr, sink, out, iters, out_list, out_cont, out_count,
                        slot);
        }
    }
}

// The first tier with a refill per group: a group that is done with its pair
// takes the next list entry at once, while the other groups of its wave go on
// with theirs, so a wave no longer waits for the longest of its pairs.  Each
// trip through the loop runs, per group, either the set-up of a new pair or one
// EPA iteration (the two are executed one after the other by the wave when its
// groups differ).  The fetch and the loop's exit test are at the top of the
// loop, where the whole wave is converged: a ballot of the idle groups' first
// lanes, one atomicAdd by lane 0 for all of them, the base read with
// readfirstlane, and each idle group's entry at its rank among them (no
// cross-lane exchange under divergence).  Everything after is group-uniform,
// and the cross-lane operations of the EPA code stay within a group.
template <int G, class W>
__global__ __launch_bounds__(64, (epa_min_waves<G, W>())) void epa_refill_kernel(const ShapeRecord *__restrict__ rec,
                                                      const float4 *__restrict__ verts, const float *__restrict__ poses,
                                                      const uint32_t *__restrict__ pairs, float prediction,
                                                      const EpaItem *__restrict__ queue, const uint32_t *in_list,
                                                      const uint32_t *in_count, uint32_t *next,
                                                      const uint32_t *in_cont, const EpaCont *cont_recs,
                                                      ContSink sink, DevContact *__restrict__ out,
                                                      uint32_t *__restrict__ iters, uint32_t *out_list,
                                                      uint32_t *out_cont, uint32_t *out_count, float *orig_scratch) {
    static_assert(G < 64, "a wave-per-pair tier has nothing to refill");
    constexpr int kGroups = 64 / G;
    __shared__ W w[kGroups];
    const uint32_t group = threadIdx.x / G;
    const uint32_t sub = __lane_id() % G;
    const uint32_t gid = blockIdx.x * kGroups + group;
    W *const wg = &w[group];
    float *orig = orig_scratch + (size_t)gid * W::kV * kOrigFloats;
    const uint32_t count = __hip_atomic_load(in_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool busy = false, drained = false;
    uint32_t slot = 0, status = EPA_OK, ei = 0, cont = kNoCont;
    EpaProblem pr{};
    Staged st{nullptr, nullptr};
    EpaRegs<W> r{Heap<W>{wg, 0}, 0, 0, 0, 0, kFltMax, 0.0f};
    f3 p1{0.0f, 0.0f, 0.0f}, p2{0.0f, 0.0f, 0.0f}, n{0.0f, 0.0f, 0.0f};
    for (;;) {
        const bool idle = !busy && !drained;  // group-uniform
        const uint64_t want = __ballot(idle && sub == 0);
        if (want != 0ull) {  // wave-uniform
            uint32_t base = 0;
            if (__lane_id() == 0) base = atomicAdd(next, (uint32_t)__popcll(want));
            base = __builtin_amdgcn_readfirstlane(base);
            const uint32_t q = base + (uint32_t)__popcll(want & ((1ull << (group * G)) - 1ull));
            if (idle && q >= count) {
                drained = true;
            } else if (idle) {
                slot = in_list ? in_list[q] : q;
                const uint32_t ci = in_cont ? in_cont[q] : kNoCont;
                pr = epa_problem(rec, verts, poses, pairs, queue + slot);
                status = epa_begin<G>(wg, orig, pr.pos, pr.a, pr.b, queue + slot,
                                      ci != kNoCont ? cont_recs + ci : nullptr, st, r, p1, p2, n, ei, cont);
                busy = true;
            }
        }
        if (__ballot(!drained) == 0ull) break;
        if (busy) {
            if (status == EPA_RUNNING)
                status = epa_step<G>(wg, orig, pr.pos, pr.a, pr.b, st, sink, r, p1, p2, n, ei, cont);
            if (status != EPA_RUNNING) {
                if (sub == 0)
                    epa_finish(rec, verts, poses, pairs, prediction, queue + slot, status, p1, p2, n, ei, cont, out,
                               iters, out_list, out_cont, out_count, slot);
                busy = false;
            }
        }
    }
}

// Last tier: one wave per pair, polytope in global memory.
__global__ __launch_bounds__(64) void epa_global_kernel(const ShapeRecord *__restrict__ rec,
                                                        const float4 *__restrict__ verts,
                                                        const float *__restrict__ poses,
                                                        const uint32_t *__restrict__ pairs, float prediction,
                                                        const EpaItem *__restrict__ queue,
                                                        DevContact *__restrict__ out, uint32_t *__restrict__ iters,
                                                        const uint32_t *in_list, const uint32_t *in_count,
                                                        uint32_t *next, EpaLarge *scratch, float *orig_scratch) {
    EpaLarge *w = scratch + blockIdx.x;
    float *orig = orig_scratch + (size_t)blockIdx.x * EpaLarge::kV * kOrigFloats;
    const uint32_t count = __hip_atomic_load(in_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        uint32_t q = 0;
        if (__lane_id() == 0) q = atomicAdd(next, 1u);
        q = __builtin_amdgcn_readfirstlane(q);
        if (q >= count) break;
        epa_item<64>(w, orig, rec, verts, poses, pairs, prediction, queue + in_list[q], nullptr, ContSink{nullptr, nullptr, 0},
                     out, iters, nullptr, nullptr, nullptr, q);
    }
}

template <int G, class W>
hipError_t launch_epa_tier(const DeviceShapes &s, const float *poses, const uint32_t *pairs, float prediction,
                           uint64_t max_items, const EpaItem *queue, const uint32_t *in_list, const uint32_t *in_count,
                           uint32_t *next, const uint32_t *in_cont, const EpaCont *cont_recs, ContSink sink,
                           DevContact *out, uint32_t *iters, uint32_t *out_list, uint32_t *out_cont,
                           uint32_t *out_count, float *orig, uint64_t max_groups, hipStream_t stream,
                           bool refill = false) {
    hipError_t e;
    int dev = 0, cus = 0, per_cu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const void *kernel = (const void *)epa_tier_kernel<G, W>;
    if constexpr (G < 64) {
        if (refill) kernel = (const void *)epa_refill_kernel<G, W>;
    }
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0)) != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    constexpr uint64_t kGroups = 64 / G;
    uint64_t waves = (uint64_t)cus * (uint64_t)per_cu;
    const uint64_t need = (max_items + kGroups - 1) / kGroups;
    if (waves > need) waves = need;
    if (waves * kGroups > max_groups) waves = max_groups / kGroups;
    if (waves < 1) waves = 1;
    if constexpr (G < 64) {
        if (refill) {
            hipLaunchKernelGGL((epa_refill_kernel<G, W>), dim3((uint32_t)waves), dim3(64), 0, stream, s.rec, s.verts,
                               poses, pairs, prediction, queue, in_list, in_count, next, in_cont, cont_recs, sink, out,
                               iters, out_list, out_cont, out_count, orig);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((epa_tier_kernel<G, W>), dim3((uint32_t)waves), dim3(64), 0, stream, s.rec, s.verts, poses,
                       pairs, prediction, queue, in_list, in_count, next, in_cont, cont_recs, sink, out, iters,
                       out_list, out_cont, out_count, orig);
    return hipGetLastError();
}

}  // namespace

// Resident EPA groups the workspace provides orig1/orig2 buffers for, per tier.
constexpr uint64_t kEpaMaxGroups = 256ull * 16 * 8;
constexpr uint64_t kEpaMediumWaves = 256ull * 8;
// orig1/orig2 rows per first-tier group: the most vertices of any first-tier polytope.
constexpr uint64_t kSmallMaxVertices = 48;
static_assert(EpaSmall::kV <= kSmallMaxVertices && EpaSmallUnstaged::kV <= kSmallMaxVertices &&
                  EpaTiny::kV <= kSmallMaxVertices && EpaSmallWide::kV <= kSmallMaxVertices &&
                  EpaSmallLarge::kV <= kSmallMaxVertices,
              "first-tier orig buffer too small");
constexpr uint64_t kEpaBigWaves = 256ull * 2;

int g_epa_lanes_per_pair = 16;
int g_epa_polytope = 1;
int g_epa_schedule = 0;
bool g_contact_phase_timing = false;
hipEvent_t g_contact_phase_events[kContactPhases + 1] = {};

namespace {
struct ContactWorkspace {
    uint32_t *counters;  // list lengths and the tiers' next entries (reset_counters_kernel)
    EpaItem *queue;
    uint32_t *lists[3];
    uint32_t *cont_list;  // per entry of lists[0]: its continuation record or kNoCont
    EpaCont *conts;
    uint32_t cont_cap;
    EpaLarge *large;
    float *orig_large, *orig_small, *orig_medium, *orig_big;
    size_t bytes;
};

// Continuation records: one per 48 pairs (overflowing pairs are about 1% of
// mixed-shape pairs), at least 256 and at most 32768.  A pair that finds them
// all taken restarts in the next tier instead.
uint64_t cont_capacity(uint64_t m) {
    uint64_t c = m / 48;
    if (c < 256) c = 256;
    if (c > 32768) c = 32768;
    return c;
}

ContactWorkspace carve_workspace(void *base, uint64_t m) {
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    char *ws = (char *)base;
    ContactWorkspace w{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = ws ? ws + off : nullptr;
        off += up(bytes);
        return p;
    };
    w.counters = (uint32_t *)take(kNumCounters * sizeof(uint32_t));
    w.queue = (EpaItem *)take(m * sizeof(EpaItem));
    for (int k = 0; k < 3; ++k) w.lists[k] = (uint32_t *)take(m * sizeof(uint32_t));
    w.cont_list = (uint32_t *)take(m * sizeof(uint32_t));
    w.cont_cap = (uint32_t)cont_capacity(m);
    w.conts = (EpaCont *)take((size_t)w.cont_cap * sizeof(EpaCont));
    w.large = (EpaLarge *)take(kEpaGlobalWaves * sizeof(EpaLarge));
    w.orig_large = (float *)take(kEpaGlobalWaves * EpaLarge::kV * kOrigFloats * sizeof(float));
    w.orig_small = (float *)take(kEpaMaxGroups * kSmallMaxVertices * kOrigFloats * sizeof(float));
    w.orig_medium = (float *)take(kEpaMediumWaves * EpaMedium::kV * kOrigFloats * sizeof(float));
    w.orig_big = (float *)take(kEpaBigWaves * EpaBig::kV * kOrigFloats * sizeof(float));
    w.bytes = off;
    return w;
}
}  // namespace

size_t contact_workspace_bytes(uint64_t num_pairs) { return carve_workspace(nullptr, num_pairs).bytes; }

namespace {
// First EPA tier: G lanes per pair with the selected small polytope; it hands
// overflowing pairs on with continuation records.
template <int G, class W>
hipError_t launch_first(const DeviceShapes &s, const float *poses, const uint32_t *pairs, float prediction, uint64_t m,
                        const ContactWorkspace &w, uint32_t *c, DevContact *o, uint32_t *iters, hipStream_t stream,
                        bool refill) {
    return launch_epa_tier<G, W>(s, poses, pairs, prediction, m, w.queue, nullptr, c, c + 4, nullptr, nullptr,
                                 ContSink{w.conts, c + 8, w.cont_cap}, o, iters, w.lists[0], w.cont_list, c + 1,
                                 w.orig_small, kEpaMaxGroups, stream, refill);
}
template <int G>
hipError_t launch_small(const DeviceShapes &s, const float *poses, const uint32_t *pairs, float prediction, uint64_t m,
                        const ContactWorkspace &w, uint32_t *c, DevContact *o, uint32_t *iters, hipStream_t stream,
                        uint32_t flags) {
    // the first tier's schedule: a refill per group, except for short EPA runs
    const bool refill = G < 64 && (g_epa_schedule == 2 || (g_epa_schedule == 0 && !(flags & kContactShortEpa)));
    switch (g_epa_polytope) {
        case 1: return launch_first<G, EpaSmallUnstaged>(s, poses, pairs, prediction, m, w, c, o, iters, stream, refill);
        case 2: return launch_first<G, EpaTiny>(s, poses, pairs, prediction, m, w, c, o, iters, stream, refill);
        case 3: return launch_first<G, EpaSmallWide>(s, poses, pairs, prediction, m, w, c, o, iters, stream, refill);
        case 4: return launch_first<G, EpaSmallLarge>(s, poses, pairs, prediction, m, w, c, o, iters, stream, refill);
        default: return launch_first<G, EpaSmall>(s, poses, pairs, prediction, m, w, c, o, iters, stream, refill);
    }
}
}  // namespace

hipError_t launch_contacts(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                           const uint64_t *m_dev, float prediction, const float *seeds, pq_contact *out,
                           uint32_t *iters, void *workspace, hipStream_t stream, uint32_t flags) {
    if (m == 0) return hipSuccess;
    const ContactWorkspace w = carve_workspace(workspace, m);
    DevContact *o = (DevContact *)out;
    const bool timed = g_contact_phase_timing;
    auto mark = [&](int k) { return timed ? hipEventRecord(g_contact_phase_events[k], stream) : hipSuccess; };
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, w.counters);
    hipError_t e = mark(0);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(contact_gjk_kernel, dim3((uint32_t)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       s.rec, s.verts, poses, pairs, m, (const unsigned long long *)m_dev, prediction, seeds, o, iters,
                       w.queue, w.counters, flags);
    e = hipGetLastError();
    if (e == hipSuccess) e = mark(1);
    if (e != hipSuccess) return e;
    uint32_t *const c = w.counters;
    switch (g_epa_lanes_per_pair) {
        case 4: e = launch_small<4>(s, poses, pairs, prediction, m, w, c, o, iters, stream, flags); break;
        case 8: e = launch_small<8>(s, poses, pairs, prediction, m, w, c, o, iters, stream, flags); break;
        case 16: e = launch_small<16>(s, poses, pairs, prediction, m, w, c, o, iters, stream, flags); break;
        case 32: e = launch_small<32>(s, poses, pairs, prediction, m, w, c, o, iters, stream, flags); break;
        default: e = launch_small<64>(s, poses, pairs, prediction, m, w, c, o, iters, stream, flags); break;
    }
    if (e == hipSuccess) e = mark(2);
    if (e != hipSuccess) return e;
    const ContSink none{nullptr, nullptr, 0};
    e = launch_epa_tier<64, EpaMedium>(s, poses, pairs, prediction, m, w.queue, w.lists[0], c + 1, c + 5, w.cont_list,
                                       w.conts, none, o, iters, w.lists[1], nullptr, c + 2, w.orig_medium,
                                       kEpaMediumWaves, stream);
    if (e == hipSuccess) e = mark(3);
    if (e != hipSuccess) return e;
    e = launch_epa_tier<64, EpaBig>(s, poses, pairs, prediction, m, w.queue, w.lists[1], c + 2, c + 6, nullptr, nullptr,
                                    none, o, iters, w.lists[2], nullptr, c + 3, w.orig_big, kEpaBigWaves, stream);
    if (e == hipSuccess) e = mark(4);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(epa_global_kernel, dim3(kEpaGlobalWaves), dim3(64), 0, stream, s.rec, s.verts, poses, pairs,
                       prediction, w.queue, o, iters, w.lists[2], c + 3, c + 7, w.large, w.orig_large);
    e = hipGetLastError();
    if (e == hipSuccess) e = mark(5);
    return e;
}

}  // namespace pq
