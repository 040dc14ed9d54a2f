// capi.cpp -- extern "C" entry points declared in include/parry_mi355x.h.
#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "pq_internal.h"

namespace pq {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    return fail(e == hipErrorOutOfMemory ? PQ_ERR_OOM : PQ_ERR_HIP, m);
}

int check_ray_tuning(const pq_ray_tuning *t, const char *who) {
    if (!t) return fail(PQ_ERR_INVALID, std::string(who) + ": null tuning");
    const uint32_t b = t->block_size, r = t->stack_regs;
    if (t->mapping > PQ_RAY_MAPPING_BINNED || t->reserved != 0 || t->chunk > (1u << 20) ||
        (b != 0 && b != 64 && b != 128 && b != 256) || (r != 0 && r != 1 && r != 2 && r != 4))
        return fail(PQ_ERR_INVALID, std::string(who) + ": invalid tuning (mapping 0-3, chunk <= 2^20, block_size "
                                                       "0/64/128/256, stack_regs 0/1/2/4, reserved 0)");
    return PQ_OK;
}

hipError_t stream_queue(hipStream_t stream, unsigned long long **queue) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, unsigned long long *> queues;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(dev, stream);
    auto it = queues.find(key);
    if (it != queues.end()) {
        *queue = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    if ((e = hipMalloc(&p, 2 * sizeof(unsigned long long))) != hipSuccess) return e;
    if ((e = hipMemset(p, 0, 2 * sizeof(unsigned long long))) != hipSuccess) {
        (void)hipFree(p);
        return e;
    }
    queues[key] = (unsigned long long *)p;
    *queue = (unsigned long long *)p;
    return hipSuccess;
}

hipError_t ray_bin_workspace(hipStream_t stream, uint64_t num_rays, uint32_t blocks, RayBinSpace *ws) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, std::pair<uint32_t *, size_t>> spaces;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t hist_words = (size_t)kRayBins * blocks + kRayBins;
    const size_t words = hist_words + 2 * (size_t)num_rays;
    std::lock_guard<std::mutex> lock(mu);
    auto &slot = spaces[std::make_pair(dev, stream)];
    if (slot.second < words) {
        if (slot.first) {
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
            (void)hipFree(slot.first);
            slot = {nullptr, 0};
        }
        void *p = nullptr;
        if ((e = hipMalloc(&p, words * sizeof(uint32_t))) != hipSuccess) return e;
        slot = {(uint32_t *)p, words};
    }
    ws->hist = slot.first;
    ws->totals = slot.first + (size_t)kRayBins * blocks;
    ws->keyrank = slot.first + hist_words;
    ws->order = ws->keyrank + num_rays;
    return hipSuccess;
}

hipError_t stream_workspace(hipStream_t stream, size_t bytes, void **ptr) {
    static std::mutex mu;
    static std::map<std::pair<int, hipStream_t>, std::pair<void *, size_t>> spaces;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    auto &slot = spaces[std::make_pair(dev, stream)];
    if (slot.second < bytes) {
        if (slot.first) {
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
            (void)hipFree(slot.first);
            slot = {nullptr, 0};
        }
        if ((e = hipMalloc(&slot.first, bytes)) != hipSuccess) return e;
        slot.second = bytes;
    }
    *ptr = slot.first;
    return hipSuccess;
}

namespace {

// LDS budget of the ray kernel's per-lane stack (256 lanes x depth-1 entries x 4 B).
constexpr uint32_t kMaxRayStack = 64;

int upload(void **dptr, const void *src, size_t bytes) {
    PQ_HIP_CHECK(hipMalloc(dptr, bytes));
    hipError_t e = hipMemcpy(*dptr, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*dptr);
        *dptr = nullptr;
        return hip_fail(e, "hipMemcpy(host->device)");
    }
    return PQ_OK;
}

}  // namespace
}  // namespace pq

using namespace pq;

extern "C" {

size_t pq_last_error(char *buf, size_t cap) {
    const std::string &m = g_last_error;
    if (buf && cap) {
        size_t k = m.size() < cap - 1 ? m.size() : cap - 1;
        std::memcpy(buf, m.data(), k);
        buf[k] = 0;
    }
    return m.size();
}

int pq_device_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(PQ_ERR_INVALID, "pq_device_malloc: null ptr");
    PQ_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
    return PQ_OK;
}

int pq_device_free(void *ptr) {
    if (ptr) PQ_HIP_CHECK(hipFree(ptr));
    return PQ_OK;
}

int pq_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return PQ_OK;
    PQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return PQ_OK;
}

int pq_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return PQ_OK;
    PQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PQ_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return PQ_OK;
}

int pq_stream_synchronize(void *stream) {
    PQ_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return PQ_OK;
}

int pq_bvh_from_leaves(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t capacity,
                       uint32_t *num_nodes, uint32_t *depth) {
    if (n && !aabbs) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves: null aabbs");
    if (!num_nodes) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves: null num_nodes");
    std::vector<LeafBox> leaves(n);
    std::memcpy(leaves.data(), aabbs, sizeof(LeafBox) * (size_t)n);
    std::vector<pq_bvh_node_wide> nodes;
    uint32_t d = 0;
    build_binned_bvh(leaves, nodes, &d);
    *num_nodes = (uint32_t)nodes.size();
    if (depth) *depth = d;
    if (capacity < nodes.size())
        return fail(PQ_ERR_CAPACITY, "pq_bvh_from_leaves: need " + std::to_string(nodes.size()) + " records");
    if (!nodes.empty()) {
        if (!out) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves: null output");
        std::memcpy(out, nodes.data(), nodes.size() * sizeof(pq_bvh_node_wide));
    }
    return PQ_OK;
}

int pq_trimesh_create(const float *vertices, uint32_t num_vertices, const uint32_t *indices, uint32_t num_triangles,
                      const pq_bvh_node_wide *nodes, uint32_t num_nodes, pq_trimesh **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_trimesh_create: null output handle");
    *out = nullptr;
    if (num_triangles == 0 || !indices) return fail(PQ_ERR_INVALID, "pq_trimesh_create: empty index buffer");
    if (!vertices || num_vertices == 0) return fail(PQ_ERR_INVALID, "pq_trimesh_create: empty vertex buffer");
    for (uint64_t k = 0; k < 3ull * num_triangles; ++k)
        if (indices[k] >= num_vertices)
            return fail(PQ_ERR_INVALID, "pq_trimesh_create: vertex index " + std::to_string(indices[k]) +
                                            " out of range at position " + std::to_string(k));

    auto *m = new pq_trimesh();
    uint32_t depth = 0;
    if (nodes) {
        bool ok = false;
        depth = bvh_depth(nodes, num_nodes, num_triangles, &ok);
        if (!ok) {
            delete m;
            return fail(PQ_ERR_INVALID, "pq_trimesh_create: the supplied Bvh is not a well-formed tree over the triangles");
        }
        m->host_nodes.assign(nodes, nodes + num_nodes);
    } else {
        std::vector<LeafBox> leaves(num_triangles);
        for (uint32_t t = 0; t < num_triangles; ++t) {
            const float *a = vertices + 3 * (size_t)indices[3 * t];
            const float *b = vertices + 3 * (size_t)indices[3 * t + 1];
            const float *c = vertices + 3 * (size_t)indices[3 * t + 2];
            // Triangle::local_aabb (aabb_triangle.rs:16-30): min(min(a,b),c).
            for (int d = 0; d < 3; ++d) {
                leaves[t].mins[d] = std::fmin(std::fmin(a[d], b[d]), c[d]);
                leaves[t].maxs[d] = std::fmax(std::fmax(a[d], b[d]), c[d]);
            }
        }
        build_binned_bvh(leaves, m->host_nodes, &depth);
    }
    if (depth > kMaxRayStack + 1) {
        delete m;
        return fail(PQ_ERR_UNSUPPORTED, "pq_trimesh_create: Bvh depth " + std::to_string(depth) +
                                            " exceeds the ray kernel's stack budget (" +
                                            std::to_string(kMaxRayStack + 1) + " levels)");
    }

    // Device copy: leaves renumbered to the order the depth-first tree visits
    // them, triangle records gathered in that order.
    std::vector<pq_bvh_node_wide> dev_nodes = m->host_nodes;
    std::vector<TriRecord> recs;
    recs.reserve(num_triangles);
    for (auto &w : dev_nodes) {
        pq_bvh_node *side[2] = {&w.left, &w.right};
        for (pq_bvh_node *c : side) {
            if (leaf_count(*c) == 1) {
                uint32_t t = c->children;
                TriRecord r{};
                std::memcpy(r.a, vertices + 3 * (size_t)indices[3 * t], 12);
                std::memcpy(r.b, vertices + 3 * (size_t)indices[3 * t + 1], 12);
                std::memcpy(r.c, vertices + 3 * (size_t)indices[3 * t + 2], 12);
                r.tri_id = t;
                c->children = (uint32_t)recs.size();
                recs.push_back(r);
            }
        }
    }
    int rc = upload(&m->d_nodes, dev_nodes.data(), dev_nodes.size() * sizeof(pq_bvh_node_wide));
    if (rc == PQ_OK) rc = upload(&m->d_tris, recs.data(), recs.size() * sizeof(TriRecord));
    if (rc != PQ_OK) {
        pq_trimesh_destroy(m);
        return rc;
    }
    m->dev.nodes = (const pq_bvh_node_wide *)m->d_nodes;
    m->dev.tris = (const TriRecord *)m->d_tris;
    m->dev.num_nodes = (uint32_t)dev_nodes.size();
    m->dev.num_triangles = num_triangles;
    m->dev.depth = depth;
    bool finite = true;
    for (size_t k = 0; k < 3 * (size_t)num_vertices && finite; ++k) finite = std::isfinite(vertices[k]);
    for (const pq_bvh_node_wide &w : dev_nodes)
        for (const pq_bvh_node *c : {&w.left, &w.right})
            for (int d = 0; d < 3; ++d) finite = finite && std::isfinite(c->mins[d]) && std::isfinite(c->maxs[d]);
    m->dev.finite = finite ? 1u : 0u;
    m->dev.tuning = pq_ray_tuning{PQ_RAY_MAPPING_STATIC, 128, 0, 256, 0, 0};
    *out = m;
    return PQ_OK;
}

int pq_trimesh_destroy(pq_trimesh *m) {
    if (!m) return PQ_OK;
    if (m->d_nodes) (void)hipFree(m->d_nodes);
    if (m->d_tris) (void)hipFree(m->d_tris);
    delete m;
    return PQ_OK;
}

int pq_trimesh_info(const pq_trimesh *m, uint32_t *num_nodes, uint32_t *depth, uint32_t *num_triangles) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_info: null mesh");
    if (num_nodes) *num_nodes = m->dev.num_nodes;
    if (depth) *depth = m->dev.depth;
    if (num_triangles) *num_triangles = m->dev.num_triangles;
    return PQ_OK;
}

int pq_trimesh_set_tuning(pq_trimesh *m, const pq_ray_tuning *t) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_set_tuning: null mesh");
    const int rc = check_ray_tuning(t, "pq_trimesh_set_tuning");
    if (rc != PQ_OK) return rc;
    m->dev.tuning = *t;
    return PQ_OK;
}

int pq_trimesh_autotune(pq_trimesh *m, const pq_ray *rays, uint64_t num_rays, float max_toi, pq_ray_hit *hits,
                        void *stream, pq_ray_tuning *chosen, float *chosen_ms) {
    if (!m || !rays || !hits || num_rays == 0) return fail(PQ_ERR_INVALID, "pq_trimesh_autotune: null argument");
    const hipStream_t s = (hipStream_t)stream;
    return autotune_ray_mapping([&]() { return launch_cast_rays(m->dev, rays, num_rays, max_toi, hits, nullptr, s); },
                                m->dev.tuning, s, chosen, chosen_ms, "pq_trimesh_autotune", true);
}

int pq_trimesh_get_tuning(const pq_trimesh *m, pq_ray_tuning *t) {
    if (!m || !t) return fail(PQ_ERR_INVALID, "pq_trimesh_get_tuning: null argument");
    *t = m->dev.tuning;
    return PQ_OK;
}

int pq_trimesh_export_bvh(const pq_trimesh *m, pq_bvh_node_wide *out, uint32_t capacity) {
    if (!m || !out) return fail(PQ_ERR_INVALID, "pq_trimesh_export_bvh: null argument");
    if (capacity < m->host_nodes.size())
        return fail(PQ_ERR_CAPACITY, "pq_trimesh_export_bvh: need " + std::to_string(m->host_nodes.size()) + " records");
    std::memcpy(out, m->host_nodes.data(), m->host_nodes.size() * sizeof(pq_bvh_node_wide));
    return PQ_OK;
}

int pq_trimesh_cast_local_rays(const pq_trimesh *m, const pq_ray *rays, uint64_t n, float max_toi, int solid,
                               pq_ray_hit *hits, void *stream) {
    (void)solid;  // TriMesh parts are triangles, which ignore `solid` (ray_triangle.rs:49-62)
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays: null mesh");
    if (n && (!rays || !hits)) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays: null buffer");
    PQ_HIP_CHECK(launch_cast_rays(m->dev, rays, n, max_toi, hits, nullptr, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_cast_local_rays_counted(const pq_trimesh *m, const pq_ray *rays, uint64_t n, float max_toi, int solid,
                                       pq_ray_hit *hits, uint32_t *counts, void *stream) {
    (void)solid;
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_counted: null mesh");
    if (n && (!rays || !hits || !counts)) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_counted: null buffer");
    PQ_HIP_CHECK(launch_cast_rays(m->dev, rays, n, max_toi, hits, counts, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_cast_local_rays_with_culling(const pq_trimesh *m, const pq_ray *rays, uint64_t n, float max_toi,
                                            uint32_t culling, pq_ray_hit *hits, void *stream) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling: null mesh");
    if (culling != PQ_CULL_BACKFACES && culling != PQ_CULL_FRONTFACES)
        return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling: culling must be PQ_CULL_BACKFACES or "
                                    "PQ_CULL_FRONTFACES");
    if (n && (!rays || !hits)) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling: null buffer");
    PQ_HIP_CHECK(launch_cast_rays_culling(m->dev, rays, n, max_toi, culling, hits, nullptr, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_cast_local_rays_with_culling_counted(const pq_trimesh *m, const pq_ray *rays, uint64_t n,
                                                    float max_toi, uint32_t culling, pq_ray_hit *hits,
                                                    uint32_t *counts, void *stream) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling_counted: null mesh");
    if (culling != PQ_CULL_BACKFACES && culling != PQ_CULL_FRONTFACES)
        return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling_counted: bad culling mode");
    if (n && (!rays || !hits || !counts))
        return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_with_culling_counted: null buffer");
    PQ_HIP_CHECK(launch_cast_rays_culling(m->dev, rays, n, max_toi, culling, hits, counts, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_project_local_points(const pq_trimesh *m, const float *points, uint64_t n, int solid,
                                    pq_point_projection *out, void *stream) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_project_local_points: null mesh");
    if (n && (!points || !out)) return fail(PQ_ERR_INVALID, "pq_trimesh_project_local_points: null buffer");
    PQ_HIP_CHECK(launch_project_points(m->dev, points, n, solid, out, nullptr, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_project_local_points_counted(const pq_trimesh *m, const float *points, uint64_t n, int solid,
                                            pq_point_projection *out, uint32_t *counts, void *stream) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_project_local_points_counted: null mesh");
    if (n && (!points || !out || !counts))
        return fail(PQ_ERR_INVALID, "pq_trimesh_project_local_points_counted: null buffer");
    PQ_HIP_CHECK(launch_project_points(m->dev, points, n, solid, out, counts, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_cast_local_rays_host(const pq_trimesh *m, const pq_ray *rays, uint64_t n, float max_toi, int solid,
                                    pq_ray_hit *hits) {
    if (!m) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_host: null mesh");
    if (n == 0) return PQ_OK;
    if (!rays || !hits) return fail(PQ_ERR_INVALID, "pq_trimesh_cast_local_rays_host: null buffer");
    void *dr = nullptr, *dh = nullptr;
    int rc = pq_device_malloc(&dr, n * sizeof(pq_ray));
    if (rc == PQ_OK) rc = pq_device_malloc(&dh, n * sizeof(pq_ray_hit));
    if (rc == PQ_OK) rc = pq_memcpy_htod(dr, rays, n * sizeof(pq_ray), nullptr);
    if (rc == PQ_OK) rc = pq_trimesh_cast_local_rays(m, (const pq_ray *)dr, n, max_toi, solid, (pq_ray_hit *)dh, nullptr);
    if (rc == PQ_OK) rc = pq_memcpy_dtoh(hits, dh, n * sizeof(pq_ray_hit), nullptr);
    pq_device_free(dr);
    pq_device_free(dh);
    return rc;
}

}  // extern "C"
