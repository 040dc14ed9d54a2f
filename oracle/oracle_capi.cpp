// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" surface of the CPU restatement in parry_oracle.hpp, loaded with
// ctypes by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Independent queries are split over a std::thread pool in dynamic chunks:
// the analogue of a rayon par_iter over rays (parry's own `parallel` feature is
// unused in src/, so the pool is this file's choice, not a reference pattern).
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "parry_oracle.hpp"
#include "parry_oracle_contact.hpp"
#include "parry_oracle_point.hpp"
#include "parry_oracle_cast.hpp"
#include "parry_oracle_manifold.hpp"

using namespace oracle;

namespace {

void parallel_for(uint64_t n, int nthreads, uint64_t chunk, const std::function<void(uint64_t, uint64_t)> &fn) {
    if (nthreads <= 1 || n <= chunk) {
        fn(0, n);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                uint64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto &th : pool) th.join();
}

template <class T> TriMesh<T> *make_mesh(const T *verts, uint32_t nv, const uint32_t *idx, uint32_t nt) {
    auto *m = new TriMesh<T>();
    m->vertices.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) m->vertices[i] = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m->indices.assign(idx, idx + 3 * (size_t)nt);
    trimesh_build(*m);
    return m;
}

template <class T>
void cast_rays(const TriMesh<T> *m, const T *rays, uint64_t n, T max_toi, int solid, uint32_t *fid, T *toi,
               T *normal, uint64_t *counters, uint32_t *per_ray_w, uint32_t *per_ray_k, int nthreads,
               int brute, int culling = 0) {
    std::atomic<uint64_t> w_sum{0}, k_sum{0};
    std::atomic<uint32_t> max_stack{0};
    parallel_for(n, nthreads, 1024, [&](uint64_t b, uint64_t e) {
        RayCounters c;
        for (uint64_t i = b; i < e; ++i) {
            Ray<T> r{v3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5])};
            RayHit<T> h{};
            uint64_t w0 = c.wide_loads, k0 = c.leaf_tests;
            bool hit = brute ? trimesh_cast_ray_brute(*m, r, max_toi, &h, culling)
                             : trimesh_cast_ray(*m, r, max_toi, solid != 0, &h, &c, culling);
            if (per_ray_w) per_ray_w[i] = (uint32_t)(c.wide_loads - w0);
            if (per_ray_k) per_ray_k[i] = (uint32_t)(c.leaf_tests - k0);
            if (hit) {
                fid[i] = h.fid;
                toi[i] = h.toi;
                normal[3 * i] = h.normal.x;
                normal[3 * i + 1] = h.normal.y;
                normal[3 * i + 2] = h.normal.z;
            } else {
                fid[i] = UINT32_MAX;
                toi[i] = T(0);
                normal[3 * i] = normal[3 * i + 1] = normal[3 * i + 2] = T(0);
            }
        }
        w_sum += c.wide_loads;
        k_sum += c.leaf_tests;
        uint32_t ms = max_stack.load();
        while (c.max_stack > ms && !max_stack.compare_exchange_weak(ms, c.max_stack)) {
        }
    });
    if (counters) {
        counters[0] = w_sum.load();
        counters[1] = k_sum.load();
        counters[2] = max_stack.load();
    }
}

template <class T>
void project_points(const TriMesh<T> *m, const T *pts, uint64_t n, int solid, T *proj, uint32_t *inside,
                    uint32_t *tri, int nthreads, int brute, uint64_t *counters) {
    std::atomic<uint64_t> w_sum{0}, k_sum{0};
    parallel_for(n, nthreads, 1024, [&](uint64_t b, uint64_t e) {
        RayCounters c;
        for (uint64_t i = b; i < e; ++i) {
            V3<T> p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            PointHit<T> h{};
            bool ok = brute ? trimesh_project_point_brute(*m, p, solid != 0, &h)
                            : trimesh_project_point(*m, p, solid != 0, &h, &c);
            proj[3 * i] = ok ? h.point.x : T(0);
            proj[3 * i + 1] = ok ? h.point.y : T(0);
            proj[3 * i + 2] = ok ? h.point.z : T(0);
            inside[i] = ok && h.inside ? 1u : 0u;
            tri[i] = ok ? h.tri : UINT32_MAX;
        }
        w_sum += c.wide_loads;
        k_sum += c.leaf_tests;
    });
    if (counters) {
        counters[0] = w_sum.load();
        counters[1] = k_sum.load();
    }
}

// 64-byte record, same field order as BvhNodeWide in f32 (bvh_tree.rs:263-266, 456-467).
struct WideF32 {
    float lmins[3];
    uint32_t lchildren;
    float lmaxs[3];
    uint32_t ldata;
    float rmins[3];
    uint32_t rchildren;
    float rmaxs[3];
    uint32_t rdata;
};
static_assert(sizeof(WideF32) == 64, "wide record is 64 bytes");

}  // namespace

extern "C" {

void *orc_trimesh_create_f32(const float *verts, uint32_t nv, const uint32_t *idx, uint32_t nt) {
    return make_mesh<float>(verts, nv, idx, nt);
}
void *orc_trimesh_create_f64(const double *verts, uint32_t nv, const uint32_t *idx, uint32_t nt) {
    return make_mesh<double>(verts, nv, idx, nt);
}
void orc_trimesh_free_f32(void *m) { delete (TriMesh<float> *)m; }
void orc_trimesh_free_f64(void *m) { delete (TriMesh<double> *)m; }

uint32_t orc_trimesh_num_nodes_f32(void *m) { return (uint32_t)((TriMesh<float> *)m)->bvh.nodes.size(); }
uint32_t orc_trimesh_depth_f32(void *m) { return bvh_depth(((TriMesh<float> *)m)->bvh); }

// Copies the Bvh wide nodes as 64-byte BvhNodeWide-layout records.
void orc_trimesh_nodes_f32(void *mp, void *out) {
    auto *m = (TriMesh<float> *)mp;
    auto *o = (WideF32 *)out;
    for (size_t i = 0; i < m->bvh.nodes.size(); ++i) {
        const auto &w = m->bvh.nodes[i];
        WideF32 r;
        for (int d = 0; d < 3; ++d) {
            r.lmins[d] = w.left.mins[d];
            r.lmaxs[d] = w.left.maxs[d];
            r.rmins[d] = w.right.mins[d];
            r.rmaxs[d] = w.right.maxs[d];
        }
        r.lchildren = w.left.children;
        r.ldata = w.left.data;
        r.rchildren = w.right.children;
        r.rdata = w.right.data;
        o[i] = r;
    }
}

// counters (nullable): [sum W, sum K, max stack depth].
void orc_cast_rays_f32(void *m, const float *rays, uint64_t n, float max_toi, int solid, uint32_t *fid, float *toi,
                       float *normal, uint64_t *counters, uint32_t *per_ray_w, uint32_t *per_ray_k, int nthreads) {
    cast_rays<float>((TriMesh<float> *)m, rays, n, max_toi, solid, fid, toi, normal, counters, per_ray_w, per_ray_k,
                     nthreads, 0);
}
void orc_cast_rays_f64(void *m, const double *rays, uint64_t n, double max_toi, int solid, uint32_t *fid,
                       double *toi, double *normal, uint64_t *counters, int nthreads) {
    cast_rays<double>((TriMesh<double> *)m, rays, n, max_toi, solid, fid, toi, normal, counters, nullptr, nullptr,
                      nthreads, 0);
}
void orc_cast_rays_culling_f32(void *m, const float *rays, uint64_t n, float max_toi, int culling, uint32_t *fid,
                               float *toi, float *normal, uint64_t *counters, int nthreads, int brute) {
    cast_rays<float>((TriMesh<float> *)m, rays, n, max_toi, 0, fid, toi, normal, counters, nullptr, nullptr, nthreads,
                     brute, culling);
}
void orc_cast_rays_culling_f64(void *m, const double *rays, uint64_t n, double max_toi, int culling, uint32_t *fid,
                               double *toi, double *normal, uint64_t *counters, int nthreads) {
    cast_rays<double>((TriMesh<double> *)m, rays, n, max_toi, 0, fid, toi, normal, counters, nullptr, nullptr,
                      nthreads, 0, culling);
}
void orc_project_points_f32(void *m, const float *pts, uint64_t n, int solid, float *proj, uint32_t *inside,
                            uint32_t *tri, uint64_t *counters, int nthreads, int brute) {
    project_points<float>((TriMesh<float> *)m, pts, n, solid, proj, inside, tri, nthreads, brute, counters);
}
void orc_project_points_f64(void *m, const double *pts, uint64_t n, int solid, double *proj, uint32_t *inside,
                            uint32_t *tri, uint64_t *counters, int nthreads, int brute) {
    project_points<double>((TriMesh<double> *)m, pts, n, solid, proj, inside, tri, nthreads, brute, counters);
}
int orc_triangle_project_point_f32(const float *abc, const float *pt, int solid, float *out) {
    bool ins = false;
    V3<float> p = triangle_project_point(v3(abc[0], abc[1], abc[2]), v3(abc[3], abc[4], abc[5]),
                                         v3(abc[6], abc[7], abc[8]), v3(pt[0], pt[1], pt[2]), solid != 0, &ins);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
    return ins ? 1 : 0;
}
void orc_cast_rays_brute_f32(void *m, const float *rays, uint64_t n, float max_toi, uint32_t *fid, float *toi,
                             float *normal, int nthreads) {
    cast_rays<float>((TriMesh<float> *)m, rays, n, max_toi, 1, fid, toi, normal, nullptr, nullptr, nullptr, nthreads,
                     1);
}

// Single-AABB slab test (Aabb::cast_local_ray, ray_aabb.rs:12-49). Returns 1 on hit.
int orc_aabb_cast_ray_f32(const float *mins, const float *maxs, const float *ray, float max_toi, int solid,
                          float *out) {
    Aabb<float> a{v3(mins[0], mins[1], mins[2]), v3(maxs[0], maxs[1], maxs[2])};
    Ray<float> r{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5])};
    return aabb_cast_local_ray(a, r, max_toi, solid != 0, out) ? 1 : 0;
}

// Single-triangle ray cast (ray_triangle.rs:49-62). Returns 1 on hit; out = [toi, nx, ny, nz], *fid = 0 front / 1 back.
int orc_triangle_cast_ray_f32(const float *abc, const float *ray, float max_toi, float *out, uint32_t *fid) {
    Ray<float> r{v3(ray[0], ray[1], ray[2]), v3(ray[3], ray[4], ray[5])};
    RayHit<float> h;
    if (!triangle_cast(v3(abc[0], abc[1], abc[2]), v3(abc[3], abc[4], abc[5]), v3(abc[6], abc[7], abc[8]), r, max_toi,
                       &h))
        return 0;
    out[0] = h.toi;
    out[1] = h.normal.x;
    out[2] = h.normal.y;
    out[3] = h.normal.z;
    *fid = h.fid;
    return 1;
}

}  // extern "C"

// ---- broadphase ------------------------------------------------------------

struct OrcBvh {
    Bvh<float> bvh;
    std::vector<LeafSlot> slots;
    uint32_t n;
};

static Aabb<float> aabb_at(const float *a, uint32_t i) {
    return Aabb<float>{v3(a[6 * i], a[6 * i + 1], a[6 * i + 2]), v3(a[6 * i + 3], a[6 * i + 4], a[6 * i + 5])};
}

template <class T>
static void polyhedra_aabbs(const T *verts, const uint32_t *offsets, uint32_t n, const T *poses, T *out, int nthreads) {
    parallel_for(n, nthreads, 256, [&](uint64_t b, uint64_t e) {
        std::vector<V3<T>> pts;
        for (uint64_t i = b; i < e; ++i) {
            pts.clear();
            for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k)
                pts.push_back(v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]));
            const T *p = poses + 7 * i;
            Iso<T> m{Quat<T>{p[0], p[1], p[2], p[3]}, v3(p[4], p[5], p[6])};
            Aabb<T> a = point_cloud_aabb(m, pts.data(), pts.size());
            T *o = out + 6 * i;
            o[0] = a.mins.x; o[1] = a.mins.y; o[2] = a.mins.z;
            o[3] = a.maxs.x; o[4] = a.maxs.y; o[5] = a.maxs.z;
        }
    });
}

extern "C" {

void orc_polyhedra_aabbs_f32(const float *verts, const uint32_t *offsets, uint32_t n, const float *poses, float *out,
                             int nthreads) {
    polyhedra_aabbs<float>(verts, offsets, n, poses, out, nthreads);
}
void orc_polyhedra_aabbs_f64(const double *verts, const uint32_t *offsets, uint32_t n, const double *poses,
                             double *out, int nthreads) {
    polyhedra_aabbs<double>(verts, offsets, n, poses, out, nthreads);
}

// Bvh::from_leaves(Binned, aabbs) over n leaf AABBs (n*6 floats).
void *orc_bvh_create_f32(const float *aabbs, uint32_t n) {
    auto *h = new OrcBvh();
    std::vector<Aabb<float>> leaves(n);
    for (uint32_t i = 0; i < n; ++i) leaves[i] = aabb_at(aabbs, i);
    h->bvh = bvh_from_leaves(leaves);
    h->slots = leaf_slots(h->bvh, n);
    h->n = n;
    return h;
}
// Bvh::from_leaves(Ploc, aabbs) (bvh_tree.rs:1891-1955, bvh_ploc_build.rs:10-93).
void *orc_bvh_create_ploc_f32(const float *aabbs, uint32_t n) {
    auto *h = new OrcBvh();
    std::vector<Aabb<float>> leaves(n);
    for (uint32_t i = 0; i < n; ++i) leaves[i] = aabb_at(aabbs, i);
    h->bvh = bvh_from_leaves_ploc(leaves);
    h->slots = leaf_slots(h->bvh, n);
    h->n = n;
    return h;
}
// The Morton keys of the PLOC build (bvh_ploc_build.rs:12-21), in leaf order.
void orc_ploc_keys_f32(const float *aabbs, uint32_t n, uint64_t *out) {
    std::vector<Node<float>> leaves;
    for (uint32_t i = 0; i < n; ++i) leaves.push_back(Node<float>::leaf(aabb_at(aabbs, i), i));
    std::vector<uint64_t> k = ploc_keys(leaves);
    for (uint32_t i = 0; i < n; ++i) out[i] = k[i];
}
uint64_t orc_morton_encode_unorm(double x, double y, double z) { return morton_encode_unorm(x, y, z); }
void orc_bvh_free_f32(void *h) { delete (OrcBvh *)h; }
uint32_t orc_bvh_num_nodes_f32(void *h) { return (uint32_t)((OrcBvh *)h)->bvh.nodes.size(); }
void orc_bvh_nodes_f32(void *hp, void *out) {
    auto *h = (OrcBvh *)hp;
    auto *o = (WideF32 *)out;
    for (size_t i = 0; i < h->bvh.nodes.size(); ++i) {
        const auto &w = h->bvh.nodes[i];
        WideF32 r;
        for (int d = 0; d < 3; ++d) {
            r.lmins[d] = w.left.mins[d];
            r.lmaxs[d] = w.left.maxs[d];
            r.rmins[d] = w.right.mins[d];
            r.rmaxs[d] = w.right.maxs[d];
        }
        r.lchildren = w.left.children;
        r.ldata = w.left.data;
        r.rchildren = w.right.children;
        r.rdata = w.right.data;
        o[i] = r;
    }
}

// One broadphase frame's tree update: insert_or_update_partially(aabb_i, i, 0.0)
// for every leaf (bvh_insert.rs:209-231) followed by refit (bvh_refit.rs:170-180).
void orc_bvh_update_refit_f32(void *hp, const float *aabbs) {
    auto *h = (OrcBvh *)hp;
    for (uint32_t i = 0; i < h->n; ++i) update_leaf(h->bvh, h->slots[i], aabb_at(aabbs, i));
    refit(h->bvh);
    h->slots = leaf_slots(h->bvh, h->n);
}

// traverse_bvtt_single_tree::<false>; pairs in callback order (2 u32 each).
// Returns 0 and sets *count to the number of pairs; fills `out` up to `cap` pairs.
void orc_bvh_self_pairs_f32(void *hp, uint32_t *out, uint64_t cap, uint64_t *count, uint64_t *wide_touched) {
    auto *h = (OrcBvh *)hp;
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    BvttStats st;
    self_pairs(h->bvh, pairs, st);
    *count = pairs.size();
    if (wide_touched) *wide_touched = st.wide_touched;
    for (uint64_t i = 0; i < pairs.size() && i < cap; ++i) {
        out[2 * i] = pairs[i].first;
        out[2 * i + 1] = pairs[i].second;
    }
}

// Every unordered pair (i < j) of closed-overlapping AABBs, by exhaustive test
// (O(n^2); the ground truth for small sets).
void orc_brute_pairs_f32(const float *aabbs, uint32_t n, uint32_t *out, uint64_t cap, uint64_t *count) {
    uint64_t c = 0;
    for (uint32_t i = 0; i < n; ++i) {
        Aabb<float> a = aabb_at(aabbs, i);
        for (uint32_t j = i + 1; j < n; ++j) {
            if (a.intersects(aabb_at(aabbs, j))) {
                if (c < cap) {
                    out[2 * c] = i;
                    out[2 * c + 1] = j;
                }
                ++c;
            }
        }
    }
    *count = c;
}

}  // extern "C"

// ---- contacts ----------------------------------------------------------------

namespace {

struct ShapeDesc {  // same 32-byte layout as pq_shape_desc
    uint32_t type, first_vertex, num_vertices, reserved;
    float params[4];
};

template <class T> std::vector<Shape<T>> make_shapes(const ShapeDesc *descs, uint32_t nshapes,
                                                     const std::vector<V3<T>> &pts) {
    std::vector<Shape<T>> shapes(nshapes);
    for (uint32_t i = 0; i < nshapes; ++i) {
        shapes[i].type = descs[i].type;
        shapes[i].pts = pts.data() + descs[i].first_vertex;
        shapes[i].n = descs[i].num_vertices;
        for (int k = 0; k < 4; ++k) shapes[i].p[k] = (T)descs[i].params[k];
    }
    return shapes;
}

template <class T>
void shape_aabbs(const ShapeDesc *descs, uint32_t nshapes, const T *verts, uint32_t nverts, const T *poses, T *out) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(descs, nshapes, pts);
    for (uint32_t i = 0; i < nshapes; ++i) {
        const T *a = poses + 7 * (size_t)i;
        Iso<T> m{Quat<T>{a[0], a[1], a[2], a[3]}, v3(a[4], a[5], a[6])};
        Aabb<T> b = shape_aabb(m, shapes[i]);
        T *o = out + 6 * (size_t)i;
        o[0] = b.mins.x; o[1] = b.mins.y; o[2] = b.mins.z;
        o[3] = b.maxs.x; o[4] = b.maxs.y; o[5] = b.maxs.z;
    }
}

template <class T>
void contact_pairs(const ShapeDesc *descs, uint32_t nshapes, const T *verts, uint32_t nverts, const T *poses,
                   const uint32_t *pairs, uint64_t m, T prediction, const T *init_dirs, T *out, int32_t *status,
                   uint32_t *iters, uint32_t *epa_exit, int nthreads) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(descs, nshapes, pts);
    parallel_for(m, nthreads, 64, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; ++k) {
            uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
            const T *a = poses + 7 * (size_t)i, *c = poses + 7 * (size_t)j;
            Iso<T> p1{Quat<T>{a[0], a[1], a[2], a[3]}, v3(a[4], a[5], a[6])};
            Iso<T> p2{Quat<T>{c[0], c[1], c[2], c[3]}, v3(c[4], c[5], c[6])};
            V3<T> seed;
            const V3<T> *sp = nullptr;
            if (init_dirs) {
                seed = v3(init_dirs[3 * k], init_dirs[3 * k + 1], init_dirs[3 * k + 2]);
                sp = &seed;
            }
            Contact<T> ct{};
            ContactStats st;
            int r = contact(p1, shapes[i], p2, shapes[j], prediction, ct, sp, &st);
            status[k] = r;
            T *o = out + 13 * k;
            if (r == 1) {
                o[0] = ct.point1.x; o[1] = ct.point1.y; o[2] = ct.point1.z;
                o[3] = ct.point2.x; o[4] = ct.point2.y; o[5] = ct.point2.z;
                o[6] = ct.normal1.x; o[7] = ct.normal1.y; o[8] = ct.normal1.z;
                o[9] = ct.normal2.x; o[10] = ct.normal2.y; o[11] = ct.normal2.z;
                o[12] = ct.dist;
            } else {
                for (int q = 0; q < 13; ++q) o[q] = T(0);
            }
            if (iters) {
                iters[2 * k] = st.gjk_iters;
                iters[2 * k + 1] = st.used_epa ? st.epa_iters + 1 : 0;
            }
            if (epa_exit) epa_exit[k] = st.used_epa ? st.epa_exit : EPA_NOT_RUN;
        }
    });
}

template <class T>
void local_supports(const ShapeDesc *desc, const T *verts, uint32_t nverts, const T *dirs, uint64_t n, T *out) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(desc, 1, pts);
    for (uint64_t k = 0; k < n; ++k) {
        V3<T> p = local_support_point(shapes[0], v3(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]));
        out[3 * k] = p.x; out[3 * k + 1] = p.y; out[3 * k + 2] = p.z;
    }
}

}  // namespace

extern "C" {

// SupportMap::local_support_point of one shape along each of n directions (n x 3).
void orc_local_support_f32(const void *desc, const float *verts, uint32_t nverts, const float *dirs, uint64_t n,
                           float *out) {
    local_supports<float>((const ShapeDesc *)desc, verts, nverts, dirs, n, out);
}
void orc_local_support_f64(const void *desc, const double *verts, uint32_t nverts, const double *dirs, uint64_t n,
                           double *out) {
    local_supports<double>((const ShapeDesc *)desc, verts, nverts, dirs, n, out);
}

// Shape::compute_aabb(pose) for every shape (n x 6).
void orc_shape_aabbs_f32(const void *descs, uint32_t n, const float *verts, uint32_t nverts, const float *poses,
                         float *out) {
    shape_aabbs<float>((const ShapeDesc *)descs, n, verts, nverts, poses, out);
}

// query::contact for every (i, j) in `pairs` (m pairs).  out: 13 per pair
// (point1, point2, normal1, normal2, dist, world frame); status: 1 contact,
// 0 None, -1 shape pair not handled; iters (nullable): GJK iterations and
// EPA iterations + 1 (0 when EPA did not run); epa_exit (nullable): how EPA
// ended (EpaExit, 0 when it did not run).  init_dirs (nullable): GJK seed
// directions in shape 1's local frame, replacing parry's default seed.
void orc_contact_pairs_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                           const float *poses, const uint32_t *pairs, uint64_t m, float prediction,
                           const float *init_dirs, float *out, int32_t *status, uint32_t *iters,
                           uint32_t *epa_exit, int nthreads) {
    contact_pairs<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, init_dirs, out,
                         status, iters, epa_exit, nthreads);
}
void orc_contact_pairs_f64(const void *descs, uint32_t nshapes, const double *verts, uint32_t nverts,
                           const double *poses, const uint32_t *pairs, uint64_t m, double prediction,
                           const double *init_dirs, double *out, int32_t *status, uint32_t *iters,
                           uint32_t *epa_exit, int nthreads) {
    contact_pairs<double>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, init_dirs,
                          out, status, iters, epa_exit, nthreads);
}

}  // extern "C"

namespace {
template <class T>
void trimesh_contact_shapes(const TriMesh<T> *mesh, const T *mesh_pose, const ShapeDesc *descs, uint32_t nshapes,
                            const T *verts, uint32_t nverts, const T *poses, T prediction, T *out, int32_t *status,
                            uint32_t *tri, uint64_t *candidates, int nthreads) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(descs, nshapes, pts);
    Iso<T> p1{Quat<T>{mesh_pose[0], mesh_pose[1], mesh_pose[2], mesh_pose[3]}, v3(mesh_pose[4], mesh_pose[5], mesh_pose[6])};
    std::atomic<uint64_t> cand{0};
    parallel_for(nshapes, nthreads, 64, [&](uint64_t b, uint64_t e) {
        uint64_t local = 0;
        for (uint64_t q = b; q < e; ++q) {
            const T *a = poses + 7 * (size_t)q;
            Iso<T> p2{Quat<T>{a[0], a[1], a[2], a[3]}, v3(a[4], a[5], a[6])};
            Contact<T> ct{};
            uint32_t t = UINT32_MAX;
            int r = trimesh_contact_shape(*mesh, p1, shapes[q], p2, prediction, ct, &t, &local);
            status[q] = r;
            tri[q] = r ? t : UINT32_MAX;
            T *o = out + 13 * q;
            if (r) {
                o[0] = ct.point1.x; o[1] = ct.point1.y; o[2] = ct.point1.z;
                o[3] = ct.point2.x; o[4] = ct.point2.y; o[5] = ct.point2.z;
                o[6] = ct.normal1.x; o[7] = ct.normal1.y; o[8] = ct.normal1.z;
                o[9] = ct.normal2.x; o[10] = ct.normal2.y; o[11] = ct.normal2.z;
                o[12] = ct.dist;
            } else {
                for (int k = 0; k < 13; ++k) o[k] = T(0);
            }
        }
        cand += local;
    });
    if (candidates) *candidates = cand.load();
}
}  // namespace

extern "C" {
void orc_trimesh_contact_shapes_f32(void *m, const float *mesh_pose, const void *descs, uint32_t nshapes,
                                    const float *verts, uint32_t nverts, const float *poses, float prediction,
                                    float *out, int32_t *status, uint32_t *tri, uint64_t *candidates, int nthreads) {
    trimesh_contact_shapes<float>((TriMesh<float> *)m, mesh_pose, (const ShapeDesc *)descs, nshapes, verts, nverts,
                                  poses, prediction, out, status, tri, candidates, nthreads);
}
void orc_trimesh_contact_shapes_f64(void *m, const double *mesh_pose, const void *descs, uint32_t nshapes,
                                    const double *verts, uint32_t nverts, const double *poses, double prediction,
                                    double *out, int32_t *status, uint32_t *tri, uint64_t *candidates, int nthreads) {
    trimesh_contact_shapes<double>((TriMesh<double> *)m, mesh_pose, (const ShapeDesc *)descs, nshapes, verts, nverts,
                                   poses, prediction, out, status, tri, candidates, nthreads);
}
}  // extern "C"

// ---- shape casts ---------------------------------------------------------------
namespace {
template <class T>
void cast_pairs(const ShapeDesc *descs, uint32_t nshapes, const T *verts, uint32_t nverts, const T *poses,
                const T *vels, const uint32_t *pairs, uint64_t m, T max_toi, T target_distance, int stop,
                int compute_geometry, T *out, uint32_t *status, uint32_t *iters, int nthreads) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(descs, nshapes, pts);
    const CastOptions<T> o{max_toi, target_distance, stop != 0, compute_geometry != 0};
    parallel_for(m, nthreads, 64, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; ++k) {
            uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
            const T *a = poses + 7 * (size_t)i, *c = poses + 7 * (size_t)j;
            Iso<T> p1{Quat<T>{a[0], a[1], a[2], a[3]}, v3(a[4], a[5], a[6])};
            Iso<T> p2{Quat<T>{c[0], c[1], c[2], c[3]}, v3(c[4], c[5], c[6])};
            V3<T> v1 = v3(vels[3 * i], vels[3 * i + 1], vels[3 * i + 2]);
            V3<T> v2 = v3(vels[3 * j], vels[3 * j + 1], vels[3 * j + 2]);
            CastHit<T> h{};
            uint32_t it = 0;
            bool ok = cast_shapes(p1, v1, shapes[i], p2, v2, shapes[j], o, h, &it);
            T *q = out + 13 * k;
            if (ok) {
                q[0] = h.toi;
                q[1] = h.witness1.x; q[2] = h.witness1.y; q[3] = h.witness1.z;
                q[4] = h.witness2.x; q[5] = h.witness2.y; q[6] = h.witness2.z;
                q[7] = h.normal1.x; q[8] = h.normal1.y; q[9] = h.normal1.z;
                q[10] = h.normal2.x; q[11] = h.normal2.y; q[12] = h.normal2.z;
                status[k] = h.status;
            } else {
                for (int r = 0; r < 13; ++r) q[r] = T(0);
                status[k] = CAST_NONE;
            }
            if (iters) iters[k] = it;
        }
    });
}
}  // namespace

extern "C" {
// query::cast_shapes for every (i, j) in `pairs` with per-shape linear
// velocities (n x 3) and one ShapeCastOptions.  out: 13 per pair (toi,
// witness1, witness2, normal1, normal2; local frames); status: 0 None, else
// ShapeCastStatus + 1; iters (nullable): minkowski_ray_cast loop passes.
void orc_cast_shapes_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts, const float *poses,
                         const float *vels, const uint32_t *pairs, uint64_t m, float max_toi, float target_distance,
                         int stop, int compute_geometry, float *out, uint32_t *status, uint32_t *iters, int nthreads) {
    cast_pairs<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, vels, pairs, m, max_toi,
                      target_distance, stop, compute_geometry, out, status, iters, nthreads);
}
void orc_cast_shapes_f64(const void *descs, uint32_t nshapes, const double *verts, uint32_t nverts,
                         const double *poses, const double *vels, const uint32_t *pairs, uint64_t m, double max_toi,
                         double target_distance, int stop, int compute_geometry, double *out, uint32_t *status,
                         uint32_t *iters, int nthreads) {
    cast_pairs<double>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, vels, pairs, m, max_toi,
                       target_distance, stop, compute_geometry, out, status, iters, nthreads);
}
}  // extern "C"

// ---- contact manifolds (pfm-pfm) ------------------------------------------------
namespace {
constexpr int kManifoldPoints = 25;
template <class T>
void manifolds_pfm(const ShapeDesc *descs, uint32_t nshapes, const T *verts, uint32_t nverts, const T *poses,
                   const uint32_t *pairs, uint64_t m, T prediction, int32_t *status, T *normals, uint32_t *counts,
                   T *points, uint32_t *fids, int nthreads, const float *face_normals = nullptr,
                   const uint32_t *face_first = nullptr, const uint32_t *face_count = nullptr, uint32_t nfaces = 0,
                   const uint32_t *face_verts = nullptr, const uint32_t *face_edges = nullptr,
                   const uint32_t *shape_faces = nullptr, bool update = false) {
    std::vector<V3<T>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]);
    std::vector<Shape<T>> shapes = make_shapes(descs, nshapes, pts);
    std::vector<PolyFace<T>> faces(nfaces);
    for (uint32_t f = 0; f < nfaces; ++f)
        faces[f] = PolyFace<T>{v3((T)face_normals[3 * f], (T)face_normals[3 * f + 1], (T)face_normals[3 * f + 2]),
                               face_first[f], face_count[f]};
    if (shape_faces) {
        for (uint32_t i = 0; i < nshapes; ++i) {
            if (shapes[i].type != CONVEX_POLYHEDRON || shape_faces[2 * i + 1] == 0) continue;
            shapes[i].faces = faces.data() + shape_faces[2 * i];
            shapes[i].nfaces = shape_faces[2 * i + 1];
            shapes[i].fverts = face_verts;
            shapes[i].fedges = face_edges;
        }
    }
    parallel_for(m, nthreads, 64, [&](uint64_t b, uint64_t e) {
        Manifold<T> mf;
        for (uint64_t k = b; k < e; ++k) {
            uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
            const T *a = poses + 7 * (size_t)i, *c = poses + 7 * (size_t)j;
            Iso<T> p1{Quat<T>{a[0], a[1], a[2], a[3]}, v3(a[4], a[5], a[6])};
            Iso<T> p2{Quat<T>{c[0], c[1], c[2], c[3]}, v3(c[4], c[5], c[6])};
            mf.points.clear();
            mf.n1 = mf.n2 = v3(T(0), T(0), T(0));
            if (update) {  // the previous frame's manifold, as the output arrays hold it
                const T *pn = normals + 6 * k;
                mf.n1 = v3(pn[0], pn[1], pn[2]);
                mf.n2 = v3(pn[3], pn[4], pn[5]);
                const T *pq = points + (size_t)k * kManifoldPoints * 7;
                const uint32_t *pf = fids + (size_t)k * kManifoldPoints * 2;
                for (uint32_t t = 0; t < counts[k] && t < (uint32_t)kManifoldPoints; ++t)
                    mf.points.push_back(TrackedContact<T>{v3(pq[7 * t], pq[7 * t + 1], pq[7 * t + 2]),
                                                          v3(pq[7 * t + 3], pq[7 * t + 4], pq[7 * t + 5]), pf[2 * t],
                                                          pf[2 * t + 1], pq[7 * t + 6]});
            }
            int r = manifold_pfm_pfm(inv_mul(p1, p2), shapes[i], shapes[j], prediction, mf);
            status[k] = r;
            T *nrm = normals + 6 * k;
            nrm[0] = mf.n1.x; nrm[1] = mf.n1.y; nrm[2] = mf.n1.z;
            nrm[3] = mf.n2.x; nrm[4] = mf.n2.y; nrm[5] = mf.n2.z;
            const size_t np = mf.points.size();
            counts[k] = (uint32_t)np;
            T *q = points + (size_t)k * kManifoldPoints * 7;
            uint32_t *f = fids + (size_t)k * kManifoldPoints * 2;
            for (size_t t = 0; t < (size_t)kManifoldPoints; ++t) {
                if (t < np) {
                    const auto &c2 = mf.points[t];
                    q[7 * t + 0] = c2.p1.x; q[7 * t + 1] = c2.p1.y; q[7 * t + 2] = c2.p1.z;
                    q[7 * t + 3] = c2.p2.x; q[7 * t + 4] = c2.p2.y; q[7 * t + 5] = c2.p2.z;
                    q[7 * t + 6] = c2.dist;
                    f[2 * t] = c2.fid1;
                    f[2 * t + 1] = c2.fid2;
                } else {
                    for (int z = 0; z < 7; ++z) q[7 * t + z] = T(0);
                    f[2 * t] = f[2 * t + 1] = 0u;
                }
            }
        }
    });
}
}  // namespace

extern "C" {
// contact_manifold_pfm_pfm on a fresh manifold for every (i, j) in `pairs`.
// status: 1 ClosestPoints, 0 NoIntersection, -1 no feature map here; normals:
// 6 per pair (local_n1, local_n2); counts: points per pair (at most 25);
// points: 25 x 7 per pair (local_p1, local_p2, dist); fids: 25 x 2 per pair.
void orc_manifolds_pfm_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                           const float *poses, const uint32_t *pairs, uint64_t m, float prediction, int32_t *status,
                           float *normals, uint32_t *counts, float *points, uint32_t *fids, int nthreads) {
    manifolds_pfm<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, status,
                         normals, counts, points, fids, nthreads);
}
// The same with ConvexPolyhedron faces: nfaces faces (normal, first, count into
// face_verts / face_edges) and 2 per shape in shape_faces (first face, count).
void orc_manifolds_pfm_faces_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                                 const float *poses, const uint32_t *pairs, uint64_t m, float prediction,
                                 int32_t *status, float *normals, uint32_t *counts, float *points, uint32_t *fids,
                                 int nthreads, const float *face_normals, const uint32_t *face_first,
                                 const uint32_t *face_count, uint32_t nfaces, const uint32_t *face_verts,
                                 const uint32_t *face_edges, const uint32_t *shape_faces) {
    manifolds_pfm<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, status,
                         normals, counts, points, fids, nthreads, face_normals, face_first, face_count, nfaces,
                         face_verts, face_edges, shape_faces);
}
// The same on the previous frame's manifolds, read from and written back to the
// output arrays (status 2: kept by try_update_contacts_eps).
void orc_manifolds_pfm_update_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                                  const float *poses, const uint32_t *pairs, uint64_t m, float prediction,
                                  int32_t *status, float *normals, uint32_t *counts, float *points, uint32_t *fids,
                                  int nthreads) {
    manifolds_pfm<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, status,
                         normals, counts, points, fids, nthreads, nullptr, nullptr, nullptr, 0, nullptr, nullptr,
                         nullptr, true);
}
// The update with ConvexPolyhedron faces (the arguments of both variants above).
void orc_manifolds_pfm_faces_update_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                                        const float *poses, const uint32_t *pairs, uint64_t m, float prediction,
                                        int32_t *status, float *normals, uint32_t *counts, float *points,
                                        uint32_t *fids, int nthreads, const float *face_normals,
                                        const uint32_t *face_first, const uint32_t *face_count, uint32_t nfaces,
                                        const uint32_t *face_verts, const uint32_t *face_edges,
                                        const uint32_t *shape_faces) {
    manifolds_pfm<float>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, status,
                         normals, counts, points, fids, nthreads, face_normals, face_first, face_count, nfaces,
                         face_verts, face_edges, shape_faces, true);
}
void orc_manifolds_pfm_f64(const void *descs, uint32_t nshapes, const double *verts, uint32_t nverts,
                           const double *poses, const uint32_t *pairs, uint64_t m, double prediction, int32_t *status,
                           double *normals, uint32_t *counts, double *points, uint32_t *fids, int nthreads) {
    manifolds_pfm<double>((const ShapeDesc *)descs, nshapes, verts, nverts, poses, pairs, m, prediction, status,
                          normals, counts, points, fids, nthreads);
}
}  // extern "C"
