// capi_batch.cpp -- batched multi-mesh ray casts, their tuner and hipGraph capture.
#include <algorithm>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "pq_internal.h"

using namespace pq;

namespace pq {

// The candidates the ray tuners time: the lane mappings over block sizes and
// register stack depths (the binned mapping where its sort buffers exist), the
// persistent mapping's chunk sizes and a capped residency, and the wave
// mapping with one and four rays per claim.
std::vector<pq_ray_tuning> ray_tuning_candidates(bool binned) {
    std::vector<pq_ray_tuning> c;
    for (uint32_t block : {64u, 128u, 256u})
        for (uint32_t regs : {0u, 1u, 2u, 4u}) c.push_back({PQ_RAY_MAPPING_STATIC, 128, 0, block, regs, 0});
    if (binned)
        for (uint32_t block : {64u, 128u, 256u})
            for (uint32_t regs : {0u, 2u}) c.push_back({PQ_RAY_MAPPING_BINNED, 128, 0, block, regs, 0});
    for (uint32_t block : {64u, 256u})
        for (uint32_t regs : {0u, 2u})
            for (uint32_t chunk : {64u, 128u, 256u}) c.push_back({PQ_RAY_MAPPING_PERSISTENT, chunk, 0, block, regs, 0});
    c.push_back({PQ_RAY_MAPPING_PERSISTENT, 128, 4, 256, 0, 0});
    for (uint32_t chunk : {1u, 4u}) c.push_back({PQ_RAY_MAPPING_WAVE, chunk, 0, 64, 0, 0});
    return c;
}

int autotune_ray_mapping(const std::function<hipError_t()> &launch, pq_ray_tuning &slot, hipStream_t s,
                         pq_ray_tuning *chosen, float *chosen_ms, const char *who, bool binned) {
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventCreate(&e2);
    const pq_ray_tuning saved = slot;
    float best_ms = 1e30f;
    pq_ray_tuning best = saved;
    for (const pq_ray_tuning &c : ray_tuning_candidates(binned)) {
        if (e != hipSuccess) break;
        slot = c;
        // One untimed-for-the-result launch (also warms the caches); a mapping
        // more than 3x slower than the best so far is not timed further.
        if ((e = hipEventRecord(e0, s)) != hipSuccess || (e = launch()) != hipSuccess ||
            (e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess)
            break;
        float warm = 0.0f;
        if ((e = hipEventElapsedTime(&warm, e0, e1)) != hipSuccess) break;
        if (warm > 3.0f * best_ms) continue;
        if ((e = hipEventRecord(e1, s)) != hipSuccess) break;
        for (int k = 0; k < 3 && e == hipSuccess; ++k) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e2, s);
        if (e == hipSuccess) e = hipEventSynchronize(e2);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e1, e2);
        if (e != hipSuccess) break;
        ms /= 3.0f;
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e2) (void)hipEventDestroy(e2);
    if (e != hipSuccess) {
        slot = saved;
        return hip_fail(e, who);
    }
    slot = best;
    if (chosen) *chosen = best;
    if (chosen_ms) *chosen_ms = best_ms;
    return PQ_OK;
}

}  // namespace pq


namespace {

int batch_alloc(pq_trimesh_batch *b, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    b->allocations.push_back(*p);
    return PQ_OK;
}

constexpr uint32_t kMaxBatchSortRanges = 1u << 16;  // 16 KiB of counts per range: at most 1 GiB

// The sort ranges of the batch's binned mapping: ceil(rays / kRayBinRange) per
// mesh, in mesh order.  False if the batch is too large for the sort (sorted
// positions are 32-bit, and the counts are capped at 1 GiB).
bool batch_sort_ranges(const pq_trimesh_batch *b, std::vector<uint32_t> *range_mesh,
                       std::vector<uint32_t> *range_first) {
    const uint32_t nm = b->dev.num_meshes;
    if (b->dev.num_rays >= (1ull << 32) || nm > 65535u) return false;  // the scan's grid has one row per mesh
    range_first->assign(nm + 1, 0u);
    range_mesh->clear();
    uint64_t total = 0;
    for (uint32_t m = 0; m < nm; ++m) {
        (*range_first)[m] = (uint32_t)total;
        const uint64_t q = (b->offsets[m + 1] - b->offsets[m] + kRayBinRange - 1) / kRayBinRange;
        total += q;
        if (total > kMaxBatchSortRanges) return false;
        range_mesh->insert(range_mesh->end(), (size_t)q, m);
    }
    (*range_first)[nm] = (uint32_t)total;
    return true;
}

// Allocates the batch's sort buffers once (outside any graph capture: the
// callers are set_tuning and the autotuner).
int ensure_batch_bins(pq_trimesh_batch *b) {
    BatchBins &bb = b->dev.bins;
    if (bb.order) return PQ_OK;
    std::vector<uint32_t> rm, rf;
    if (!batch_sort_ranges(b, &rm, &rf))
        return fail(PQ_ERR_INVALID,
                    "the binned mapping: batch too large for the ray sort (2^32 rays, 65535 meshes, 65536 sort ranges)");
    const uint32_t nr = (uint32_t)rm.size();
    const uint64_t n = b->dev.num_rays;
    void *range_mesh = nullptr, *range_first = nullptr, *hist = nullptr, *first = nullptr, *keyrank = nullptr,
         *order = nullptr;
    int rc = batch_alloc(b, &range_mesh, rm.size() * sizeof(uint32_t));
    if (rc == PQ_OK) rc = batch_alloc(b, &range_first, rf.size() * sizeof(uint32_t));
    if (rc == PQ_OK) rc = batch_alloc(b, &hist, (size_t)nr * kRayBins * sizeof(uint32_t));
    if (rc == PQ_OK) rc = batch_alloc(b, &first, (size_t)b->dev.num_meshes * kRayBins * sizeof(uint32_t));
    if (rc == PQ_OK) rc = batch_alloc(b, &keyrank, n * sizeof(uint32_t));
    if (rc == PQ_OK) rc = batch_alloc(b, &order, n * sizeof(uint32_t));
    if (rc != PQ_OK) return rc;
    hipError_t e = rm.empty() ? hipSuccess : hipMemcpy(range_mesh, rm.data(), rm.size() * sizeof(uint32_t),
                                                        hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(range_first, rf.data(), rf.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "the binned mapping: upload of the sort ranges");
    bb.num_ranges = nr;
    bb.range_mesh = (const uint32_t *)range_mesh;
    bb.range_first = (const uint32_t *)range_first;
    bb.hist = (uint32_t *)hist;
    bb.first = (uint32_t *)first;
    bb.keyrank = (uint32_t *)keyrank;
    bb.order = (uint32_t *)order;
    return PQ_OK;
}

}  // namespace

extern "C" {

int pq_trimesh_batch_create(const pq_trimesh *const *meshes, uint32_t num_meshes, const uint64_t *ray_offsets,
                            pq_trimesh_batch **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_create: null output");
    *out = nullptr;
    if (num_meshes == 0 || !meshes || !ray_offsets)
        return fail(PQ_ERR_INVALID, "pq_trimesh_batch_create: empty batch");
    if (ray_offsets[0] != 0) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_create: ray_offsets[0] must be 0");
    std::vector<BatchMesh> bm(num_meshes);
    uint32_t max_depth = 1;
    for (uint32_t i = 0; i < num_meshes; ++i) {
        if (!meshes[i]) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_create: null mesh " + std::to_string(i));
        if (ray_offsets[i + 1] < ray_offsets[i])
            return fail(PQ_ERR_INVALID, "pq_trimesh_batch_create: ray_offsets must not decrease");
        const DeviceMesh &m = meshes[i]->dev;
        // A root whose right child has no leaf holds the mesh's single triangle.
        const bool single = (meshes[i]->host_nodes[0].right.data & 0x3fffffffu) == 0u;
        bm[i] = BatchMesh{m.nodes, m.tris, m.num_triangles, single ? 1u : 0u, m.num_nodes, m.finite};
        max_depth = std::max(max_depth, m.depth);
    }
    auto *b = new pq_trimesh_batch();
    b->offsets.assign(ray_offsets, ray_offsets + num_meshes + 1);
    DeviceBatch &d = b->dev;
    d.num_meshes = num_meshes;
    d.max_depth = max_depth;
    d.num_rays = ray_offsets[num_meshes];
    d.tuning = pq_ray_tuning{PQ_RAY_MAPPING_STATIC, 128, 0, 256, 0, 0};
    int rc = batch_alloc(b, (void **)&d.meshes, num_meshes * sizeof(BatchMesh));
    if (rc == PQ_OK) rc = batch_alloc(b, (void **)&d.offsets, (num_meshes + 1) * sizeof(unsigned long long));
    if (rc == PQ_OK) rc = batch_alloc(b, (void **)&d.queue, 2 * sizeof(unsigned long long));
    if (rc == PQ_OK) {
        hipError_t e = hipMemcpy((void *)d.meshes, bm.data(), num_meshes * sizeof(BatchMesh), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy((void *)d.offsets, ray_offsets, (num_meshes + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(d.queue, 0, 2 * sizeof(unsigned long long));
        if (e != hipSuccess) rc = hip_fail(e, "pq_trimesh_batch_create upload");
    }
    if (rc != PQ_OK) {
        pq_trimesh_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PQ_OK;
}

int pq_trimesh_batch_destroy(pq_trimesh_batch *b) {
    if (!b) return PQ_OK;
    for (void *p : b->allocations) (void)hipFree(p);
    delete b;
    return PQ_OK;
}

int pq_trimesh_batch_set_tuning(pq_trimesh_batch *b, const pq_ray_tuning *t) {
    if (!b) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_set_tuning: null batch");
    int rc = check_ray_tuning(t, "pq_trimesh_batch_set_tuning");
    if (rc != PQ_OK) return rc;
    if (t->mapping == PQ_RAY_MAPPING_BINNED && (rc = ensure_batch_bins(b)) != PQ_OK) return rc;
    b->dev.tuning = *t;
    return PQ_OK;
}

int pq_trimesh_batch_get_tuning(const pq_trimesh_batch *b, pq_ray_tuning *t) {
    if (!b || !t) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_get_tuning: null argument");
    *t = b->dev.tuning;
    return PQ_OK;
}

int pq_trimesh_batch_cast_local_rays(const pq_trimesh_batch *b, const pq_ray *rays, float max_toi, int solid,
                                     pq_ray_hit *hits, void *stream) {
    (void)solid;
    if (!b) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_cast_local_rays: null batch");
    if (b->dev.num_rays && (!rays || !hits)) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_cast_local_rays: null buffer");
    PQ_HIP_CHECK(launch_cast_rays_batch(b->dev, rays, max_toi, hits, nullptr, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_batch_cast_local_rays_counted(const pq_trimesh_batch *b, const pq_ray *rays, float max_toi, int solid,
                                             pq_ray_hit *hits, uint32_t *counts, void *stream) {
    (void)solid;
    if (!b) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_cast_local_rays_counted: null batch");
    if (b->dev.num_rays && (!rays || !hits || !counts))
        return fail(PQ_ERR_INVALID, "pq_trimesh_batch_cast_local_rays_counted: null buffer");
    PQ_HIP_CHECK(launch_cast_rays_batch(b->dev, rays, max_toi, hits, counts, (hipStream_t)stream));
    return PQ_OK;
}

int pq_trimesh_batch_autotune(pq_trimesh_batch *b, const pq_ray *rays, float max_toi, pq_ray_hit *hits, void *stream,
                              pq_ray_tuning *chosen, float *chosen_ms) {
    if (!b || !rays || !hits) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_autotune: null argument");
    const hipStream_t s = (hipStream_t)stream;
    // The binned mapping is a candidate when the batch fits the sort.
    std::vector<uint32_t> rm, rf;
    const bool binned = b->dev.num_rays > 0 && batch_sort_ranges(b, &rm, &rf);
    if (binned) {
        const int rc = ensure_batch_bins(b);
        if (rc != PQ_OK) return rc;
    }
    return autotune_ray_mapping(
        [&]() { return launch_cast_rays_batch(b->dev, rays, max_toi, hits, nullptr, s); }, b->dev.tuning, s, chosen,
        chosen_ms, "pq_trimesh_batch_autotune", binned);
}

int pq_trimesh_batch_capture(const pq_trimesh_batch *b, const pq_ray *rays, float max_toi, int solid,
                             pq_ray_hit *hits, void *stream, pq_graph **out) {
    (void)solid;
    if (!b || !out || !rays || !hits) return fail(PQ_ERR_INVALID, "pq_trimesh_batch_capture: null argument");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    hipStream_t own = nullptr;
    if (!s) {
        PQ_HIP_CHECK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        s = own;
    }
    auto *g = new pq_graph();
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = launch_cast_rays_batch(b->dev, rays, max_toi, hits, nullptr, s);
        e = hipStreamEndCapture(s, &g->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(own);
    if (e != hipSuccess) {
        pq_graph_destroy(g);
        return hip_fail(e, "pq_trimesh_batch_capture");
    }
    *out = g;
    return PQ_OK;
}

int pq_graph_launch(pq_graph *g, void *stream) {
    if (!g || !g->exec) return fail(PQ_ERR_INVALID, "pq_graph_launch: null graph");
    PQ_HIP_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return PQ_OK;
}

int pq_graph_destroy(pq_graph *g) {
    if (!g) return PQ_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return PQ_OK;
}

}  // extern "C"
