// cast.hip -- linear shape casts (query::cast_shapes, shape_cast.rs:274-286) on gfx950.
//
// Per pair: pos12 = pos1^-1 pos2 and vel12 = pos1^-1 (vel2 - vel1), then
// DefaultQueryDispatcher::cast_shapes (default_query_dispatcher.rs:434-540):
//   * Ball-Ball: cast_shapes_ball_ball (shape_cast_ball_ball.rs:10-68), a ray
//     against the ball of radius r1 + r2 + target_distance (ray_ball.rs:33-77);
//   * every other pair of the shape set (all support maps):
//     cast_shapes_support_map_support_map (shape_cast_support_map_support_map.rs:11-72)
//     -- gjk::directional_distance, i.e. GJK's ray cast on the Minkowski
//     difference (minkowski_ray_cast, gjk.rs:660-795), shape 1 rounded by the
//     target distance when it is positive (RoundShapeRef, round_shape.rs:317-343).
// A support-map hit with toi < 1e-5 whose options ask for the impact geometry
// needs contact_support_map_support_map(pos12, g1, g2, Real::MAX): those pairs
// are queued (wave-aggregated) and run through the batched contact kernels in
// their local-frame support-map mode, and a last kernel applies the options
// to the contacts.
#include "contact_gjk.h"
#include "pq_internal.h"

namespace pq {
namespace {

constexpr int kBlock = 256;

struct DevCastHit {
    float toi;
    float w1[3], w2[3], n1[3], n2[3];
    uint32_t status;
};
static_assert(sizeof(DevCastHit) == sizeof(pq_shape_cast_hit), "pq_shape_cast_hit layout");

// Set on a record whose impact geometry the contact step still has to supply.
constexpr uint32_t kCastPending = 0xFFFFFFF0u;

__device__ __forceinline__ void store_hit(DevCastHit *o, float toi, f3 w1, f3 w2, f3 n1, f3 n2, uint32_t status) {
    o->toi = toi;
    o->w1[0] = w1.x; o->w1[1] = w1.y; o->w1[2] = w1.z;
    o->w2[0] = w2.x; o->w2[1] = w2.y; o->w2[2] = w2.z;
    o->n1[0] = n1.x; o->n1[1] = n1.y; o->n1[2] = n1.z;
    o->n2[0] = n2.x; o->n2[1] = n2.y; o->n2[2] = n2.z;
    o->status = status;
}
__device__ __forceinline__ void store_none(DevCastHit *o) {
    const f3 z{0.0f, 0.0f, 0.0f};
    store_hit(o, 0.0f, z, z, z, z, PQ_CAST_NONE);
}

// SupportMap::local_support_point_toward (support_map.rs:302-304) for a unit
// direction: Ball and Capsule use it as given (ball.rs:247-249,
// capsule.rs:412-418); every other shape is its local_support_point.
__device__ __forceinline__ f3 local_support_toward(const ShapeRef &s, f3 u) {
    if (s.type == PQ_SHAPE_BALL) return mul3(u, s.p0);
    if (s.type == PQ_SHAPE_CAPSULE) {
        const f3 a{0.0f, -s.p0, 0.0f}, b{0.0f, s.p0, 0.0f};
        return dot3(u, a) > dot3(u, b) ? add3(a, mul3(u, s.p1)) : add3(b, mul3(u, s.p1));
    }
    return local_support(s, u);
}

// CSOPoint::from_shapes (cso_point.rs:69-78) with shape 1 rounded by `border`
// when border > 0: the direction normalized, shape 1's support toward it plus
// it times the border (round_shape.rs:330-343).
__device__ __forceinline__ Cso cso_rounded(const iso &pos12, const ShapeRef &g1, float border, const ShapeRef &g2,
                                           f3 dir) {
    f3 a;
    if (border > 0.0f) {
        const f3 u = div3(dir, sqrtf(norm_sq3(dir)));
        a = add3(local_support_toward(g1, u), mul3(u, border));
    } else {
        a = local_support(g1, dir);
    }
    const f3 b = iso_point(pos12, local_support(g2, qrot_inv(pos12.q, neg3(dir))));
    return Cso{sub3(a, b), a, b};
}

__device__ __forceinline__ Cso cso_shift(const Cso &c, f3 d) { return Cso{add3(c.p, d), c.o1, c.o2}; }

// minkowski_ray_cast (gjk.rs:660-795): the ray from the origin along ray_dir
// against shape 1 - pos12 * shape 2, max_time_of_impact = Real::MAX, without
// the fixed-point feature.  False for None; `passes` counts the loop's passes.
__device__ bool minkowski_ray_cast(const iso &pos12, const ShapeRef &g1, float border, const ShapeRef &g2, f3 ray_dir,
                                   Simplex &s, float &toi, f3 &normal, uint32_t &passes) {
    const float eps_rel = sqrtf(kEpsTol);
    const float ray_length = sqrtf(norm_sq3(ray_dir));
    passes = 0;
    if (rel_eq1(ray_length, 0.0f)) return false;
    float ltoi = 0.0f;
    f3 origin{0.0f, 0.0f, 0.0f};
    const f3 cdir = div3(ray_dir, ray_length);
    f3 dir = neg3(cdir);
    f3 ldir = dir;
    sx_reset(s, cso_shift(cso_rounded(pos12, g1, border, g2, dir), neg3(origin)));
    f3 proj = sx_project(s);
    float max_bound = kFltMax;
    bool last_chance = false;
    for (uint32_t niter = 0;;) {
        const float old_max_bound = max_bound;
        {
            const f3 mp = neg3(proj);
            const float sq = norm_sq3(mp);  // Unit::try_new_and_get(-proj, eps_tol)
            if (!(sq > kEpsTol * kEpsTol)) {
                toi = ltoi / ray_length;
                normal = ldir;
                return true;
            }
            const float nrm = sqrtf(sq);
            dir = div3(mp, nrm);
            max_bound = nrm;
        }
        Cso sp;
        if (max_bound >= old_max_bound) {
            last_chance = true;  // CSOPoint::single_point (cso_point.rs:42-44)
            const f3 p = add3(proj, origin);
            sp = Cso{p, p, f3{0.0f, 0.0f, 0.0f}};
        } else {
            sp = cso_rounded(pos12, g1, border, g2, dir);
        }
        if (last_chance && ltoi > 0.0f) {
            toi = ltoi / ray_length;
            normal = ldir;
            return true;
        }
        // ray_toi_with_halfspace(support point, dir, current ray) (ray_halfspace.rs:9-39)
        bool have_t = false;
        float t = 0.0f;
        {
            const f3 dpos = sub3(sp.p, origin);
            const float denom = dot3(dir, cdir);
            if (!rel_eq1(denom, 0.0f)) {
                t = dot3(dir, dpos) / denom;
                have_t = t >= 0.0f;
            }
        }
        if (have_t) {
            if (dot3(dir, cdir) < 0.0f && t > 0.0f) {  // a new lower bound: move the ray's origin
                ldir = dir;
                ltoi += t;
                if (ltoi / ray_length > kFltMax) return false;
                const f3 shift = mul3(cdir, t);
                origin = add3(origin, shift);
                max_bound = kFltMax;
                // VoronoiSimplex::modify_pnts: the points of slots 0..=dim
                const f3 back = neg3(shift);
                s.v[0].p = add3(s.v[0].p, back);
                if (s.dim >= 1) s.v[1].p = add3(s.v[1].p, back);
                if (s.dim >= 2) s.v[2].p = add3(s.v[2].p, back);
                if (s.dim >= 3) s.v[3].p = add3(s.v[3].p, back);
                last_chance = false;
            }
        } else if (dot3(dir, cdir) > kEpsTol) {
            return false;  // the ray leaves the support half-space: a miss
        }
        if (last_chance) return false;
        const float min_bound = -dot3(dir, sub3(sp.p, origin));
        if (max_bound - min_bound <= eps_rel * max_bound) return false;
        sx_add(s, cso_shift(sp, neg3(origin)));
        proj = sx_project(s);
        if (s.dim == 3) {
            if (min_bound >= kEpsTol) return false;
            toi = ltoi / ray_length;
            normal = ldir;
            return true;  // the origin is inside the Minkowski difference
        }
        ++niter;
        passes = niter;
        if (niter == kGjkMaxIters) return false;
    }
}

__device__ __forceinline__ f3 load3(const float *p, size_t i) { return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

__global__ __launch_bounds__(kBlock) void cast_kernel(const ShapeRecord *__restrict__ rec, const float4 *__restrict__ verts,
                                                      const float *__restrict__ poses, const float *__restrict__ vels,
                                                      const uint32_t *__restrict__ pairs, uint64_t m,
                                                      pq_shape_cast_options o, DevCastHit *__restrict__ out,
                                                      uint32_t *__restrict__ iters, uint32_t *__restrict__ list,
                                                      uint32_t *__restrict__ sub_pairs,
                                                      unsigned long long *__restrict__ pending) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool pend = false;
    uint32_t i = 0, j = 0;
    if (k < m) {
        i = pairs[2 * k];
        j = pairs[2 * k + 1];
        const iso pos1 = load_iso(poses + 7 * (size_t)i), pos2 = load_iso(poses + 7 * (size_t)j);
        const iso pos12 = iso_inv_mul(pos1, pos2);
        const f3 vel12 = qrot_inv(pos1.q, sub3(load3(vels, j), load3(vels, i)));
        const ShapeRef g1 = load_shape(rec, verts, i), g2 = load_shape(rec, verts, j);
        DevCastHit *const h = out + k;
        uint32_t passes = 0;
        if (g1.type == PQ_SHAPE_BALL && g2.type == PQ_SHAPE_BALL) {
            // cast_shapes_ball_ball with ray_toi_with_ball(center, r1 + r2 + target, ray, solid = true)
            const float rsum = g1.p0 + g2.p0 + o.target_distance;
            const float radius = rsum;
            const f3 center = neg3(pos12.t);
            const f3 origin{0.0f, 0.0f, 0.0f};
            const f3 dcenter = sub3(origin, center);
            const float a = norm_sq3(vel12);
            const float b = dot3(dcenter, vel12);
            const float c = norm_sq3(dcenter) - radius * radius;
            bool hit = true, inside = false;
            float toi = 0.0f;
            if (a == 0.0f) {
                if (c > 0.0f) hit = false;
                else inside = true;
            } else if (c > 0.0f && b > 0.0f) {
                hit = false;
            } else {
                const float delta = b * b - a * c;
                if (delta < 0.0f) {
                    hit = false;
                } else {
                    const float t = (-b - sqrtf(delta)) / a;
                    if (t <= 0.0f) inside = true;
                    else toi = t;
                }
            }
            if (hit && toi > o.max_time_of_impact) hit = false;
            if (hit) {
                const f3 dpt = sub3(add3(origin, mul3(vel12, toi)), center);
                f3 n1, n2, w1, w2;
                if (radius == 0.0f) {
                    n1 = f3{1.0f, 0.0f, 0.0f};
                    n2 = qrot_inv(pos12.q, neg3(f3{1.0f, 0.0f, 0.0f}));
                    w1 = w2 = f3{0.0f, 0.0f, 0.0f};
                } else {
                    n1 = div3(dpt, radius);
                    n2 = qrot_inv(pos12.q, neg3(n1));
                    w1 = mul3(n1, g1.p0);
                    w2 = mul3(n2, g2.p0);
                }
                if (!o.stop_at_penetration && toi < 1.0e-5f && dot3(n1, vel12) >= 0.0f) {
                    store_none(h);
                } else {
                    const bool pen = inside && norm_sq3(center) < rsum * rsum;
                    store_hit(h, toi, w1, w2, n1, n2, pen ? PQ_CAST_PENETRATING : PQ_CAST_CONVERGED);
                }
            } else {
                store_none(h);
            }
        } else {
            Simplex s;
            float toi = 0.0f;
            f3 n1{0.0f, 0.0f, 0.0f};
            const float border = o.target_distance > 0.0f ? o.target_distance : 0.0f;
            if (!minkowski_ray_cast(pos12, g1, border, g2, vel12, s, toi, n1, passes)) {
                store_none(h);
            } else {
                f3 w1{0.0f, 0.0f, 0.0f}, w2{0.0f, 0.0f, 0.0f};
                if (toi != 0.0f) sx_points(s, s.dim == 3, w1, w2);  // gjk::result (gjk.rs:797-818)
                if (toi > o.max_time_of_impact) {
                    store_none(h);
                } else if ((o.compute_impact_geometry_on_penetration || !o.stop_at_penetration) && toi < 1.0e-5f) {
                    h->toi = toi;
                    h->status = kCastPending;
                    pend = true;
                } else {
                    store_hit(h, toi, sub3(w1, mul3(n1, o.target_distance)),
                              qrot_inv(pos12.q, sub3(w2, pos12.t)), n1, qrot_inv(pos12.q, neg3(n1)),
                              toi == 0.0f ? PQ_CAST_PENETRATING : PQ_CAST_CONVERGED);
                }
            }
        }
        if (iters) iters[k] = passes;
    }
    // Wave-aggregated append of the pairs that need the contact.
    const uint64_t mask = __ballot(pend);
    if (mask) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(pending, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        if (pend) {
            const uint64_t p = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            list[p] = (uint32_t)k;
            sub_pairs[2 * p] = i;
            sub_pairs[2 * p + 1] = j;
        }
    }
}

// The impact geometry of the queued pairs (shape_cast_support_map_support_map.rs:41-61).
__global__ __launch_bounds__(kBlock) void cast_finish_kernel(const float *__restrict__ poses,
                                                             const float *__restrict__ vels,
                                                             const uint32_t *__restrict__ sub_pairs,
                                                             const uint32_t *__restrict__ list,
                                                             const unsigned long long *__restrict__ pending,
                                                             const pq_contact *__restrict__ contacts,
                                                             pq_shape_cast_options o, DevCastHit *__restrict__ out) {
    const uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= *pending) return;
    const uint32_t k = list[p], i = sub_pairs[2 * p], j = sub_pairs[2 * p + 1];
    DevCastHit *const h = out + k;
    const pq_contact c = contacts[p];
    if (c.has_contact != 1u) {
        store_none(h);
        return;
    }
    const iso pos1 = load_iso(poses + 7 * (size_t)i);
    const f3 vel12 = qrot_inv(pos1.q, sub3(load3(vels, j), load3(vels, i)));
    const f3 n1{c.normal1[0], c.normal1[1], c.normal1[2]};
    const float normal_vel = dot3(n1, vel12);
    if (!o.stop_at_penetration && normal_vel >= 0.0f) {
        store_none(h);
        return;
    }
    store_hit(h, h->toi, f3{c.point1[0], c.point1[1], c.point1[2]}, f3{c.point2[0], c.point2[1], c.point2[2]}, n1,
              f3{c.normal2[0], c.normal2[1], c.normal2[2]}, PQ_CAST_PENETRATING);
}

__global__ void zero_count_kernel(unsigned long long *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

struct CastWorkspace {
    uint32_t *list, *sub_pairs;
    unsigned long long *pending;
    pq_contact *contacts;
    void *contact_ws;
    size_t bytes;
};

CastWorkspace carve(void *base, uint64_t m) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    char *ws = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = ws ? ws + off : nullptr;
        off += up(bytes);
        return p;
    };
    CastWorkspace w{};
    w.list = (uint32_t *)take(m * sizeof(uint32_t));
    w.sub_pairs = (uint32_t *)take(2 * m * sizeof(uint32_t));
    w.pending = (unsigned long long *)take(sizeof(unsigned long long));
    w.contacts = (pq_contact *)take(m * sizeof(pq_contact));
    w.contact_ws = take(contact_workspace_bytes(m));
    w.bytes = off;
    return w;
}

}  // namespace

size_t cast_workspace_bytes(uint64_t m) { return carve(nullptr, m ? m : 1).bytes; }

hipError_t launch_casts(const DeviceShapes &s, const float *poses, const float *vels, const uint32_t *pairs, uint64_t m,
                        const pq_shape_cast_options &o, pq_shape_cast_hit *out, uint32_t *iters, void *workspace,
                        hipStream_t stream) {
    if (m == 0) return hipSuccess;
    const CastWorkspace w = carve(workspace, m);
    const dim3 grid((uint32_t)((m + kBlock - 1) / kBlock)), block(kBlock);
    DevCastHit *const h = (DevCastHit *)out;
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, stream, w.pending);
    hipLaunchKernelGGL(cast_kernel, grid, block, 0, stream, s.rec, s.verts, poses, vels, pairs, m, o, h, iters, w.list,
                       w.sub_pairs, w.pending);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_contacts(s, poses, w.sub_pairs, m, (const uint64_t *)w.pending, kFltMax, nullptr, w.contacts, nullptr,
                        w.contact_ws, stream, kContactLocalSupportMaps);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cast_finish_kernel, grid, block, 0, stream, poses, vels, w.sub_pairs, w.list, w.pending,
                       w.contacts, o, h);
    return hipGetLastError();
}

}  // namespace pq
