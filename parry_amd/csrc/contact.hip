// contact.hip -- batched query::contact on shape pairs (gfx950).
//
// Stage 1 (one lane per pair): pos12 = pos1.inv_mul(pos2), the dispatch of
// DefaultQueryDispatcher::contact (src/query/default_query_dispatcher.rs:302-356)
// for the shapes held here, then
//   * Ball-Ball: the closed form of contact_ball_ball.rs:9-42;
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

// Queued EPA work item: pair index and GJK's final simplex.
struct EpaItem {
    uint32_t pair;
    uint32_t dim;
    uint32_t pad[2];
    float v[4][9];  // point, orig1, orig2
};

__device__ __forceinline__ void finish_contact(const iso &pos1, const iso &pos2, const iso &pos12, f3 p1, f3 p2_1,
                                               f3 n1, DevContact &c) {
    // contact_support_map_support_map (contact_support_map_support_map.rs:18-26)
    const float dist = dot3(sub3(p2_1, p1), n1);
    const f3 p2 = qrot_inv(pos12.q, sub3(p2_1, pos12.t));
    const f3 n2 = qrot_inv(pos12.q, neg3(n1));
    // Contact::transform_by_mut
    const f3 w1 = iso_point(pos1, p1), w2 = iso_point(pos2, p2);
    const f3 m1 = qrot(pos1.q, n1), m2 = qrot(pos2.q, n2);
    c.p1[0] = w1.x; c.p1[1] = w1.y; c.p1[2] = w1.z;
    c.p2[0] = w2.x; c.p2[1] = w2.y; c.p2[2] = w2.z;
    c.n1[0] = m1.x; c.n1[1] = m1.y; c.n1[2] = m1.z;
    c.n2[0] = m2.x; c.n2[1] = m2.y; c.n2[2] = m2.z;
    c.dist = dist;
    c.has = 1;
}

__device__ __forceinline__ DevContact no_contact(uint32_t has) {
    DevContact c;
    for (int k = 0; k < 3; ++k) c.p1[k] = c.p2[k] = c.n1[k] = c.n2[k] = 0.0f;
    c.dist = 0.0f;
    c.has = has;
    return c;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__global__ void reset_counters_kernel(uint32_t *counters) {
    if (threadIdx.x < 4) counters[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// Stage 1
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void contact_gjk_kernel(
    const ShapeRecord *__restrict__ rec, const float4 *__restrict__ verts, const float *__restrict__ poses,
    const uint32_t *__restrict__ pairs, uint64_t m, const unsigned long long *__restrict__ m_dev, float prediction,
    const float *__restrict__ seeds, DevContact *__restrict__ out, uint32_t *__restrict__ iters,
    EpaItem *__restrict__ queue, uint32_t *__restrict__ counters) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (m_dev) {
        const uint64_t md = *m_dev;
        if (md < m) m = md;
    }
    bool need_epa = false;
    Simplex s;
    uint32_t pair_id = 0;
    if (k < m) {
        pair_id = (uint32_t)k;
        const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
        const iso pos12 = iso_inv_mul(pos1, pos2);
        const ShapeRef g1 = load_shape(rec, verts, i), g2 = load_shape(rec, verts, j);
        DevContact c = no_contact(0);
        uint32_t gi = 0;
        if (g1.type == PQ_SHAPE_BALL && g2.type == PQ_SHAPE_BALL) {
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
                const f3 w1 = iso_point(pos1, a), w2 = iso_point(pos2, b);
                const f3 m1 = qrot(pos1.q, n1), m2 = qrot(pos2.q, n2);
                c.p1[0] = w1.x; c.p1[1] = w1.y; c.p1[2] = w1.z;
                c.p2[0] = w2.x; c.p2[1] = w2.y; c.p2[2] = w2.z;
                c.n1[0] = m1.x; c.n1[1] = m1.y; c.n1[2] = m1.z;
                c.n2[0] = m2.x; c.n2[1] = m2.y; c.n2[2] = m2.z;
                c.dist = sqrtf(d2) - sum;
                c.has = 1;
            }
            out[k] = c;
        } else if (g1.type == PQ_SHAPE_BALL || g2.type == PQ_SHAPE_BALL) {
            out[k] = no_contact(kUnsupportedPair);
        } else {
            f3 dir;
            if (seeds) {
                dir = f3{seeds[3 * k], seeds[3 * k + 1], seeds[3 * k + 2]};
            } else {
                const float sq = norm_sq3(pos12.t);
                dir = sq > kEps * kEps ? div3(pos12.t, sqrtf(sq)) : f3{1.0f, 0.0f, 0.0f};
            }
            sx_reset(s, cso_from_shapes(pos12, g1, g2, dir));
            const GjkOut r = gjk_closest_points(pos12, g1, g2, prediction, s);
            gi = r.iters;
            if (r.kind == GJK_CLOSEST) {
                finish_contact(pos1, pos2, pos12, r.p1, r.p2, r.n, c);
                out[k] = c;
            } else if (r.kind == GJK_NO_INTERSECTION) {
                out[k] = c;
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
// Stage 2/3: EPA, one wave per pair
// ---------------------------------------------------------------------------
struct EpaFace {
    uint16_t pts[3];
    uint16_t adj[3];
    float n[3];
    float bc[3];
    uint32_t deleted;
};

template <int V, int F, int H, int S, int K> struct EpaState {
    static constexpr int kV = V, kF = F, kH = H, kS = S, kK = K;
    float vert[V][9];
    EpaFace face[F];
    uint32_t heap_id[H];
    float heap_nd[H];
    uint32_t sil[S];  // face | opp << 16
    uint32_t stk[K];  // face | opp << 16
};
// LDS variant: room for parry's 100-iteration cap in vertices; faces, heap and
// stacks sized for the common case (overflow re-runs in global memory).
using EpaSmall = EpaState<112, 224, 224, 64, 192>;
using EpaLarge = EpaState<112, 8192, 8192, 2048, 8192>;

__device__ __forceinline__ f3 wave_support(const ShapeRef &s, f3 d) {
    if (s.type != PQ_SHAPE_CONVEX_POLYHEDRON) return local_support(s, d);
    // First strict maximum of p.d over the vertex list, as a wave reduction:
    // if vertex 0's product is NaN nothing ever beats it; otherwise the result
    // is the lowest index among the largest non-NaN products.
    const float4 p0 = s.v[0];
    const float d0 = dot3(f3{p0.x, p0.y, p0.z}, d);
    if (d0 != d0) return f3{p0.x, p0.y, p0.z};
    const uint32_t lane = __lane_id();
    float best = -INFINITY;
    uint32_t bi = 0xffffffffu;
    for (uint32_t k = lane; k < s.n; k += 64) {
        const float4 p = s.v[k];
        const float dd = dot3(f3{p.x, p.y, p.z}, d);
        if (dd == dd && (bi == 0xffffffffu || dd > best)) {
            best = dd;
            bi = k;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const uint32_t oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
    }
    const float4 p = s.v[bi];
    return f3{p.x, p.y, p.z};
}

__device__ __forceinline__ Cso wave_cso(const iso &pos12, const ShapeRef &g1, const ShapeRef &g2, f3 dir) {
    const f3 a = wave_support(g1, dir);
    const f3 b = iso_point(pos12, wave_support(g2, qrot_inv(pos12.q, neg3(dir))));
    return Cso{sub3(a, b), a, b};
}

template <class W> __device__ __forceinline__ f3 vpt(const W *w, uint32_t i) {
    return f3{w->vert[i][0], w->vert[i][1], w->vert[i][2]};
}
template <class W> __device__ __forceinline__ void vset(W *w, uint32_t i, const Cso &c) {
    w->vert[i][0] = c.p.x; w->vert[i][1] = c.p.y; w->vert[i][2] = c.p.z;
    w->vert[i][3] = c.o1.x; w->vert[i][4] = c.o1.y; w->vert[i][5] = c.o1.z;
    w->vert[i][6] = c.o2.x; w->vert[i][7] = c.o2.y; w->vert[i][8] = c.o2.z;
}

// BinaryHeap<FaceId> (max-heap on neg_dist; equal keys compare Equal).
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
        w->heap_id[pos] = eid;
        w->heap_nd[pos] = end;
        return pos;
    }
    __device__ bool push(uint32_t id, float nd) {
        if (len >= (uint32_t)W::kH) return false;
        w->heap_id[len] = id;
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
        w->heap_id[pos] = eid;
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
            w->heap_id[0] = iid;
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

enum : uint32_t { EPA_OK = 0, EPA_NONE = 1, EPA_OVERFLOW = 2 };

// Face::new (epa3.rs:96-118): fills face f; returns proj_inside.
template <class W>
__device__ bool face_new(W *w, uint32_t f, uint32_t a, uint32_t b, uint32_t c, uint32_t j0, uint32_t j1, uint32_t j2) {
    EpaFace &F = w->face[f];
    F.pts[0] = (uint16_t)a; F.pts[1] = (uint16_t)b; F.pts[2] = (uint16_t)c;
    F.adj[0] = (uint16_t)j0; F.adj[1] = (uint16_t)j1; F.adj[2] = (uint16_t)j2;
    F.deleted = 0;
    const f3 pa = vpt(w, a), pb = vpt(w, b), pc = vpt(w, c);
    Loc l;
    bool inside;
    const f3 proj = tri_origin(pa, pb, pc, l, inside);
    float bc[3] = {0.0f, 0.0f, 0.0f};
    bool in;
    if (l.kind == LOC_VERTEX) {
        bc[l.id] = 1.0f;
        const float e = kEps * 100.0f;
        in = inside || norm_sq3(proj) < e * e;
    } else if (l.kind == LOC_EDGE) {
        const uint32_t i0 = l.id == 1 ? 1u : 0u, i1 = l.id == 0 ? 1u : 2u;
        bc[i0] = l.b0;
        bc[i1] = l.b1;
        const float e = kEps * 100.0f;
        in = inside || norm_sq3(proj) < e * e;
    } else if (l.kind == LOC_FACE) {
        bc[0] = l.b0; bc[1] = l.b1; bc[2] = l.b2;
        in = true;
    } else {
        in = false;
    }
    F.bc[0] = bc[0]; F.bc[1] = bc[1]; F.bc[2] = bc[2];
    // ccw_face_normal (utils/ccw_face_normal.rs:19-27), zero if degenerate
    const f3 r = cross3(sub3(pb, pa), sub3(pc, pa));
    const float sq = norm_sq3(r);
    f3 n = f3{0.0f, 0.0f, 0.0f};
    if (sq > kEps * kEps) n = div3(r, sqrtf(sq));
    F.n[0] = n.x; F.n[1] = n.y; F.n[2] = n.z;
    return in;
}

template <class W> __device__ __forceinline__ f3 fnormal(const W *w, uint32_t f) {
    return f3{w->face[f].n[0], w->face[f].n[1], w->face[f].n[2]};
}

template <class W> __device__ void face_points(const W *w, uint32_t f, f3 &p1, f3 &p2) {
    const EpaFace &F = w->face[f];
    f3 o1[3], o2[3];
    for (int q = 0; q < 3; ++q) {
        const uint32_t v = F.pts[q];
        o1[q] = f3{w->vert[v][3], w->vert[v][4], w->vert[v][5]};
        o2[q] = f3{w->vert[v][6], w->vert[v][7], w->vert[v][8]};
    }
    p1 = add3(add3(mul3(o1[0], F.bc[0]), mul3(o1[1], F.bc[1])), mul3(o1[2], F.bc[2]));
    p2 = add3(add3(mul3(o2[0], F.bc[0]), mul3(o2[1], F.bc[1])), mul3(o2[2], F.bc[2]));
}

template <class W> __device__ __forceinline__ uint32_t next_ccw(const W *w, uint32_t f, uint32_t id) {
    if (w->face[f].pts[0] == id) return 1;
    if (w->face[f].pts[1] == id) return 2;
    return 0;
}

// Face::can_be_seen_by (epa3.rs:147-163).
template <class W> __device__ bool seen_by(const W *w, uint32_t f, uint32_t point, uint32_t opp) {
    const EpaFace &F = w->face[f];
    const f3 p0 = vpt(w, F.pts[opp]);
    const f3 p1 = vpt(w, F.pts[(opp + 1) % 3]);
    const f3 p2 = vpt(w, F.pts[(opp + 2) % 3]);
    const f3 pt = vpt(w, point);
    if (dot3(sub3(pt, p0), fnormal(w, f)) >= -kEpsTol) return true;
    const float E = kEps * 100.0f;
    return norm_sq3(cross3(sub3(p2, p1), sub3(pt, p1))) <= E * E;  // Triangle::is_affinely_dependent
}

// EPA::closest_points (epa3.rs:428-651) for one pair, executed by a whole wave.
template <class W>
__device__ uint32_t epa_run(W *w, const iso &pos12, const ShapeRef &g1, const ShapeRef &g2, const EpaItem &it,
                            f3 &p1, f3 &p2, f3 &n, uint32_t &epa_iters) {
    const float eps_tol = kEps * 100.0f;
    uint32_t nv = 0, nf = 0;
    Heap<W> heap{w, 0};
    epa_iters = 0;
    for (uint32_t i = 0; i <= it.dim; ++i) {
        for (int q = 0; q < 9; ++q) w->vert[i][q] = it.v[i][q];
    }
    nv = it.dim + 1;
    if (it.dim == 0) {
        p1 = p2 = f3{0.0f, 0.0f, 0.0f};
        n = f3{0.0f, 1.0f, 0.0f};
        return EPA_OK;
    } else if (it.dim == 3) {
        const f3 dp1 = sub3(vpt(w, 1), vpt(w, 0));
        const f3 dp2 = sub3(vpt(w, 2), vpt(w, 0));
        const f3 dp3 = sub3(vpt(w, 3), vpt(w, 0));
        if (dot3(cross3(dp1, dp2), dp3) > 0.0f) {
            for (int q = 0; q < 9; ++q) {
                const float t = w->vert[1][q];
                w->vert[1][q] = w->vert[2][q];
                w->vert[2][q] = t;
            }
        }
        const bool in0 = face_new(w, 0, 0, 1, 2, 3, 1, 2);
        const bool in1 = face_new(w, 1, 1, 3, 2, 3, 2, 0);
        const bool in2 = face_new(w, 2, 0, 2, 3, 0, 1, 3);
        const bool in3 = face_new(w, 3, 0, 3, 1, 2, 1, 0);
        nf = 4;
        const bool ins[4] = {in0, in1, in2, in3};
        for (uint32_t f = 0; f < 4; ++f) {
            if (ins[f]) {
                const float dist = dot3(fnormal(w, f), vpt(w, f));
                if (-dist > kEpsTol) return EPA_NONE;  // FaceId::new -> None
                heap.push(f, -dist);
            }
        }
        if (!(in0 || in1 || in2 || in3)) return EPA_NONE;
    } else {
        if (it.dim == 1) {
            // orthonormal_subspace_basis([dpt]): first vector a x dpt
            const f3 dpt = sub3(vpt(w, 1), vpt(w, 0));
            f3 a = fabsf(dpt.x) > fabsf(dpt.y) ? f3{dpt.z, 0.0f, -dpt.x} : f3{0.0f, -dpt.z, dpt.y};
            a = div3(a, sqrtf(norm_sq3(a)));
            vset(w, 2, wave_cso(pos12, g1, g2, cross3(a, dpt)));
            nv = 3;
        }
        face_new(w, 0, 0, 1, 2, 1, 1, 1);
        face_new(w, 1, 0, 2, 1, 0, 0, 0);
        nf = 2;
        heap.push(0, 0.0f);
        heap.push(1, 0.0f);
    }
    if (heap.len == 0) return EPA_NONE;
    uint32_t best_id = w->heap_id[0];
    float max_dist = kFltMax;
    float old_dist = 0.0f;
    uint32_t niter = 0;
    uint32_t fid;
    float fnd;
    while (heap.pop(fid, fnd)) {
        if (w->face[fid].deleted) continue;
        const EpaFace face = w->face[fid];
        const f3 fn = f3{face.n[0], face.n[1], face.n[2]};
        const Cso c = wave_cso(pos12, g1, g2, fn);
        if (nv >= (uint32_t)W::kV) return EPA_OVERFLOW;
        const uint32_t sp = nv;
        vset(w, sp, c);
        ++nv;
        const float cand = dot3(c.p, fn);
        if (cand < max_dist) {
            best_id = fid;
            max_dist = cand;
        }
        const float curr_dist = -fnd;
        if (max_dist - curr_dist < eps_tol || (fabsf(curr_dist - old_dist) < kEps && cand < max_dist)) {
            face_points(w, best_id, p1, p2);
            n = fnormal(w, best_id);
            epa_iters = niter + 1;
            return EPA_OK;
        }
        old_dist = curr_dist;
        w->face[fid].deleted = 1;
        // compute_silhouette for the three neighbours, depth first (explicit stack).
        uint32_t ns = 0, sk = 0;
        for (int q = 2; q >= 0; --q) {
            const uint32_t adj = face.adj[q];
            w->stk[sk++] = adj | (next_ccw(w, adj, face.pts[q]) << 16);
        }
        while (sk > 0) {
            const uint32_t e = w->stk[--sk];
            const uint32_t id = e & 0xffffu, opp = e >> 16;
            if (w->face[id].deleted) continue;
            if (!seen_by(w, id, sp, opp)) {
                if (ns >= (uint32_t)W::kS) return EPA_OVERFLOW;
                w->sil[ns++] = e;
            } else {
                w->face[id].deleted = 1;
                const uint32_t a1 = (opp + 2) % 3, a2 = opp;
                const uint32_t adj1 = w->face[id].adj[a1], adj2 = w->face[id].adj[a2];
                const uint32_t o1 = next_ccw(w, adj1, w->face[id].pts[a1]);
                const uint32_t o2 = next_ccw(w, adj2, w->face[id].pts[a2]);
                if (sk + 2 > (uint32_t)W::kK) return EPA_OVERFLOW;
                w->stk[sk++] = adj2 | (o2 << 16);
                w->stk[sk++] = adj1 | (o1 << 16);
            }
        }
        const uint32_t first_new = nf;
        if (ns == 0) return EPA_NONE;
        for (uint32_t q = 0; q < ns; ++q) {
            const uint32_t e = w->sil[q];
            const uint32_t ef = e & 0xffffu, opp = e >> 16;
            if (w->face[ef].deleted) continue;
            if (nf >= (uint32_t)W::kF) return EPA_OVERFLOW;
            const uint32_t new_id = nf;
            const uint32_t pt1 = w->face[ef].pts[(opp + 2) % 3];
            const uint32_t pt2 = w->face[ef].pts[(opp + 1) % 3];
            const bool in = face_new(w, new_id, pt1, pt2, sp, ef, new_id + 1, new_id - 1);
            w->face[ef].adj[(opp + 1) % 3] = (uint16_t)new_id;
            ++nf;
            if (in) {
                const float dist = dot3(fnormal(w, new_id), vpt(w, w->face[new_id].pts[0]));
                if (dist < curr_dist) {
                    face_points(w, fid, p1, p2);  // `face` is unchanged: its points and bcoords
                    n = fn;
                    epa_iters = niter + 1;
                    return EPA_OK;
                }
                if (-dist > kEpsTol) return EPA_NONE;
                if (!heap.push(new_id, -dist)) return EPA_OVERFLOW;
            }
        }
        if (first_new == nf) return EPA_NONE;
        w->face[first_new].adj[2] = (uint16_t)(nf - 1);
        w->face[nf - 1].adj[1] = (uint16_t)first_new;
        ++niter;
        if (niter > 100) break;
    }
    epa_iters = niter + 1;
    face_points(w, best_id, p1, p2);
    n = fnormal(w, best_id);
    return EPA_OK;
}

template <class W>
__device__ void epa_item(W *w, const ShapeRecord *rec, const float4 *verts, const float *poses,
                         const uint32_t *pairs, const EpaItem &it, DevContact *out, uint32_t *iters,
                         uint32_t *overflow_list, uint32_t *counters, uint32_t slot) {
    const uint32_t k = it.pair;
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
    const iso pos12 = iso_inv_mul(pos1, pos2);
    const ShapeRef g1 = load_shape(rec, verts, i), g2 = load_shape(rec, verts, j);
    f3 p1, p2, n;
    uint32_t ei = 0;
    const uint32_t st = epa_run(w, pos12, g1, g2, it, p1, p2, n, ei);
    if (__lane_id() != 0) return;
    if (st == EPA_OVERFLOW) {
        if (overflow_list) {
            overflow_list[atomicAdd(&counters[1], 1u)] = slot;
            return;
        }
        out[k] = no_contact(kUnsupportedPair);  // cannot happen with the large buffers
        return;
    }
    if (st == EPA_OK) {
        DevContact c;
        finish_contact(pos1, pos2, pos12, p1, p2, n, c);
        out[k] = c;
    } else {
        out[k] = no_contact(0);  // EPA failed: NoIntersection (contact_support_map_support_map.rs:75-76)
    }
    if (iters) iters[2 * k + 1] = ei;
}

__global__ __launch_bounds__(64) void epa_lds_kernel(const ShapeRecord *__restrict__ rec,
                                                     const float4 *__restrict__ verts, const float *__restrict__ poses,
                                                     const uint32_t *__restrict__ pairs,
                                                     const EpaItem *__restrict__ queue, DevContact *__restrict__ out,
                                                     uint32_t *__restrict__ iters, uint32_t *overflow_list,
                                                     uint32_t *counters) {
    __shared__ EpaSmall w;
    const uint32_t count = __hip_atomic_load(&counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
        const EpaItem it = queue[q];
        epa_item(&w, rec, verts, poses, pairs, it, out, iters, overflow_list, counters, q);
    }
}

__global__ __launch_bounds__(64) void epa_global_kernel(const ShapeRecord *__restrict__ rec,
                                                        const float4 *__restrict__ verts,
                                                        const float *__restrict__ poses,
                                                        const uint32_t *__restrict__ pairs,
                                                        const EpaItem *__restrict__ queue,
                                                        DevContact *__restrict__ out, uint32_t *__restrict__ iters,
                                                        const uint32_t *overflow_list, uint32_t *counters,
                                                        EpaLarge *scratch) {
    EpaLarge *w = scratch + blockIdx.x;
    const uint32_t count = __hip_atomic_load(&counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
        const EpaItem it = queue[overflow_list[q]];
        epa_item(w, rec, verts, poses, pairs, it, out, iters, nullptr, counters, q);
    }
}

}  // namespace

size_t contact_workspace_bytes(uint64_t num_pairs) {
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    return up(16) + up(num_pairs * sizeof(EpaItem)) + up(num_pairs * sizeof(uint32_t)) +
           up(kEpaGlobalWaves * sizeof(EpaLarge));
}

hipError_t launch_contacts(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                           const uint64_t *m_dev, float prediction, const float *seeds, pq_contact *out,
                           uint32_t *iters, void *workspace, hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    char *ws = (char *)workspace;
    uint32_t *counters = (uint32_t *)ws;
    EpaItem *queue = (EpaItem *)(ws + up(16));
    uint32_t *overflow = (uint32_t *)(ws + up(16) + up(m * sizeof(EpaItem)));
    EpaLarge *scratch = (EpaLarge *)(ws + up(16) + up(m * sizeof(EpaItem)) + up(m * sizeof(uint32_t)));
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters);
    hipLaunchKernelGGL(contact_gjk_kernel, dim3((uint32_t)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream,
                       s.rec, s.verts, poses, pairs, m, (const unsigned long long *)m_dev, prediction, seeds,
                       (DevContact *)out, iters, queue, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    uint64_t waves = (uint64_t)cus * 8;
    if (waves > m) waves = m;
    hipLaunchKernelGGL(epa_lds_kernel, dim3((uint32_t)waves), dim3(64), 0, stream, s.rec, s.verts, poses, pairs,
                       queue, (DevContact *)out, iters, overflow, counters);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(epa_global_kernel, dim3(kEpaGlobalWaves), dim3(64), 0, stream, s.rec, s.verts, poses, pairs,
                       queue, (DevContact *)out, iters, overflow, counters, scratch);
    return hipGetLastError();
}

}  // namespace pq
