// capi_broadphase.cpp -- extern "C" entry points for shapes and the broadphase Bvh.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pq_internal.h"

using namespace pq;

namespace {

int dev_alloc(pq_bvh *b, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    b->allocations.push_back(*p);
    return PQ_OK;
}

constexpr uint32_t kMaxBroadphaseDepth = 65;

}  // namespace

extern "C" {

int pq_shapes_create(const pq_shape_desc *shapes, uint32_t n, const float *vertices, uint32_t num_vertices,
                     pq_shapes **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_shapes_create: null output");
    *out = nullptr;
    if (n && !shapes) return fail(PQ_ERR_INVALID, "pq_shapes_create: null shapes");
    if (num_vertices && !vertices) return fail(PQ_ERR_INVALID, "pq_shapes_create: null vertices");
    std::vector<ShapeRecord> rec(n);
    for (uint32_t i = 0; i < n; ++i) {
        const pq_shape_desc &d = shapes[i];
        if (d.reserved != 0) return fail(PQ_ERR_INVALID, "pq_shapes_create: reserved field must be 0");
        switch (d.type) {
            case PQ_SHAPE_BALL:
            case PQ_SHAPE_CUBOID:
            case PQ_SHAPE_CAPSULE:
            case PQ_SHAPE_CONE:
                break;
            case PQ_SHAPE_CONVEX_POLYHEDRON:
                if (d.num_vertices == 0 || (uint64_t)d.first_vertex + d.num_vertices > num_vertices)
                    return fail(PQ_ERR_INVALID, "pq_shapes_create: shape " + std::to_string(i) +
                                                    " has an empty or out-of-range vertex range");
                // params[0..1] of a polyhedron's record hold its face range, written only by
                // pq_shapes_set_polyhedron_faces together with the face tables.
                if (d.params[0] != 0.0f || d.params[1] != 0.0f || d.params[2] != 0.0f || d.params[3] != 0.0f)
                    return fail(PQ_ERR_INVALID, "pq_shapes_create: ConvexPolyhedron " + std::to_string(i) +
                                                    " must have zero params");
                break;
            case PQ_SHAPE_TRIANGLE:
                if (d.num_vertices != 3 || (uint64_t)d.first_vertex + 3 > num_vertices)
                    return fail(PQ_ERR_INVALID, "pq_shapes_create: triangle " + std::to_string(i) +
                                                    " needs 3 vertices in range");
                break;
            default:
                return fail(PQ_ERR_INVALID, "pq_shapes_create: unknown shape type " + std::to_string(d.type));
        }
        rec[i].type = d.type;
        rec[i].first_vertex = d.first_vertex;
        rec[i].num_vertices = d.num_vertices;
        rec[i].pad = 0;
        std::memcpy(rec[i].param, d.params, sizeof(rec[i].param));
    }
    std::vector<float4> v4(num_vertices);
    for (uint32_t k = 0; k < num_vertices; ++k)
        v4[k] = make_float4(vertices[3 * k], vertices[3 * k + 1], vertices[3 * k + 2], 0.0f);
    auto *s = new pq_shapes();
    s->host_rec = rec;
    hipError_t e = hipMalloc(&s->d_rec, rec.size() * sizeof(ShapeRecord) + 16);
    if (e == hipSuccess) e = hipMalloc(&s->d_verts, v4.size() * sizeof(float4) + 16);
    if (e == hipSuccess && n) e = hipMemcpy(s->d_rec, rec.data(), rec.size() * sizeof(ShapeRecord), hipMemcpyHostToDevice);
    if (e == hipSuccess && num_vertices)
        e = hipMemcpy(s->d_verts, v4.data(), v4.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        pq_shapes_destroy(s);
        return hip_fail(e, "pq_shapes_create upload");
    }
    s->dev = DeviceShapes{(const ShapeRecord *)s->d_rec, (const float4 *)s->d_verts, n, num_vertices};
    *out = s;
    return PQ_OK;
}

int pq_shapes_destroy(pq_shapes *s) {
    if (!s) return PQ_OK;
    if (s->d_rec) (void)hipFree(s->d_rec);
    if (s->d_verts) (void)hipFree(s->d_verts);
    if (s->d_faces) (void)hipFree(s->d_faces);
    delete s;
    return PQ_OK;
}

int pq_shapes_set_polyhedron_faces(pq_shapes *s, const pq_polyhedron_face *faces, uint32_t num_faces,
                                   const uint32_t *face_vertices, const uint32_t *face_edges, uint32_t num_entries,
                                   const uint32_t *shape_faces) {
    if (!s || !shape_faces) return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: null shapes or ranges");
    if ((num_faces && !faces) || (num_entries && (!face_vertices || !face_edges)))
        return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: null face tables");
    const uint32_t n = s->dev.n;
    std::vector<ShapeRecord> rec = s->host_rec;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t first = shape_faces[2 * i], count = shape_faces[2 * i + 1];
        if (rec[i].type != PQ_SHAPE_CONVEX_POLYHEDRON) {
            if (count != 0)
                return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: shape " + std::to_string(i) +
                                                " is not a ConvexPolyhedron");
            continue;
        }
        if ((uint64_t)first + count > num_faces)
            return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: shape " + std::to_string(i) +
                                            " has an out-of-range face range");
        for (uint32_t f = first; f < first + count; ++f) {
            const pq_polyhedron_face &F = faces[f];
            if (F.count < 3 || (uint64_t)F.first + F.count > num_entries)
                return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: face " + std::to_string(f) +
                                                " has fewer than 3 vertices or an out-of-range list");
            for (uint32_t k = F.first; k < F.first + F.count; ++k)
                if (face_vertices[k] >= rec[i].num_vertices)
                    return fail(PQ_ERR_INVALID, "pq_shapes_set_polyhedron_faces: face " + std::to_string(f) +
                                                    " names a vertex the polyhedron does not have");
        }
        std::memcpy(&rec[i].param[0], &first, 4);
        std::memcpy(&rec[i].param[1], &count, 4);
    }
    const size_t fb = ((size_t)num_faces * sizeof(pq_polyhedron_face) + 255) / 256 * 256;
    const size_t eb = ((size_t)num_entries * sizeof(uint32_t) + 255) / 256 * 256;
    void *buf = nullptr;
    hipError_t e = hipMalloc(&buf, fb + 2 * eb + 256);
    if (e == hipSuccess && num_faces) e = hipMemcpy(buf, faces, num_faces * sizeof(pq_polyhedron_face), hipMemcpyHostToDevice);
    if (e == hipSuccess && num_entries)
        e = hipMemcpy((char *)buf + fb, face_vertices, num_entries * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && num_entries)
        e = hipMemcpy((char *)buf + fb + eb, face_edges, num_entries * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(s->d_rec, rec.data(), n * sizeof(ShapeRecord), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (buf) (void)hipFree(buf);
        return hip_fail(e, "pq_shapes_set_polyhedron_faces");
    }
    if (s->d_faces) (void)hipFree(s->d_faces);
    s->d_faces = buf;
    s->host_rec = rec;
    s->dev.faces = (const pq_polyhedron_face *)buf;
    s->dev.face_verts = (const uint32_t *)((char *)buf + fb);
    s->dev.face_edges = (const uint32_t *)((char *)buf + fb + eb);
    return PQ_OK;
}

int pq_shapes_compute_aabbs(const pq_shapes *s, const pq_isometry *poses, float *aabbs, void *stream) {
    if (!s) return fail(PQ_ERR_INVALID, "pq_shapes_compute_aabbs: null shapes");
    if (s->dev.n && (!poses || !aabbs)) return fail(PQ_ERR_INVALID, "pq_shapes_compute_aabbs: null buffer");
    PQ_HIP_CHECK(launch_shape_aabbs(s->dev, (const float *)poses, aabbs, (hipStream_t)stream));
    return PQ_OK;
}

namespace {
// The broadphase handle of a depth-first tree (host copy `nodes`) over n leaves:
// leaf and parent slots, and the refit schedule (non-root wide nodes by height).
int bvh_handle_from_nodes(const std::vector<pq_bvh_node_wide> &nodes, uint32_t n, uint32_t depth, pq_bvh **out) {
    if (depth > kMaxBroadphaseDepth)
        return fail(PQ_ERR_UNSUPPORTED, "Bvh depth " + std::to_string(depth) + " exceeds " +
                                            std::to_string(kMaxBroadphaseDepth));
    const uint32_t nn = (uint32_t)nodes.size();
    std::vector<uint32_t> leaf_slot(n, 0), parent_slot(nn ? nn : 1, UINT32_MAX);
    for (uint32_t w = 0; w < nn; ++w) {
        const pq_bvh_node *side[2] = {&nodes[w].left, &nodes[w].right};
        for (uint32_t s = 0; s < 2; ++s) {
            const pq_bvh_node &c = *side[s];
            const uint32_t lc = leaf_count(c);
            if (lc == 1) leaf_slot[c.children] = 2 * w + s;
            else if (lc > 1) parent_slot[c.children] = 2 * w + s;
        }
    }
    // Heights above the leaves: depth-first order puts children after parents, so
    // one reverse sweep suffices.  The root's own box is not stored, so it is not
    // refitted; trees with at most two leaves have no other wide node.
    std::vector<uint32_t> height(nn, 0);
    uint32_t max_h = 0;
    for (uint32_t w = nn; w-- > 0;) {
        uint32_t h = 1;
        const pq_bvh_node *side[2] = {&nodes[w].left, &nodes[w].right};
        for (const pq_bvh_node *c : side)
            if (leaf_count(*c) > 1) h = std::max(h, height[c->children] + 1);
        height[w] = h;
        if (w != 0) max_h = std::max(max_h, h);
    }
    // offs[h] = number of nodes of height <= h: level h occupies [offs[h-1], offs[h]).
    std::vector<uint32_t> offs(max_h + 1, 0), level_nodes(nn > 0 ? nn - 1 : 0);
    for (uint32_t w = 1; w < nn; ++w) ++offs[height[w]];
    for (uint32_t h = 1; h <= max_h; ++h) offs[h] += offs[h - 1];
    {
        std::vector<uint32_t> fill(offs);  // next free position of each level: starts at offs[h-1]
        for (uint32_t h = max_h; h >= 1; --h) fill[h] = offs[h - 1];
        for (uint32_t w = 1; w < nn; ++w) level_nodes[fill[height[w]]++] = w;
    }
    auto *b = new pq_bvh();
    DeviceBvh &d = b->dev;
    b->level_offsets = offs;
    d.num_levels = max_h;
    d.level_offsets_host = b->level_offsets.data();
    d.num_leaves = n;
    d.num_nodes = nn;
    d.depth = depth;
    int rc = PQ_OK;
    size_t scan = 0;
    hipError_t e = bvh_scan_temp_bytes(n, &scan);
    if (e != hipSuccess) rc = hip_fail(e, "rocprim::exclusive_scan (size query)");
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.nodes, nn * sizeof(pq_bvh_node_wide));
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.leaf_slot, n * sizeof(uint32_t));
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.parent_slot, parent_slot.size() * sizeof(uint32_t));
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.pair_counts, ((size_t)n + 1) * sizeof(uint32_t));
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.pair_offsets, ((size_t)n + 1) * sizeof(unsigned long long));
    if (rc == PQ_OK) rc = dev_alloc(b, &d.scan_temp, scan);
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.level_nodes, level_nodes.size() * sizeof(uint32_t));
    if (rc == PQ_OK) rc = dev_alloc(b, (void **)&d.level_offsets, offs.size() * sizeof(uint32_t));
    d.scan_temp_bytes = scan;
    if (rc == PQ_OK) {
        e = hipSuccess;
        if (nn) e = hipMemcpy(d.nodes, nodes.data(), nn * sizeof(pq_bvh_node_wide), hipMemcpyHostToDevice);
        if (e == hipSuccess && n)
            e = hipMemcpy(d.leaf_slot, leaf_slot.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d.parent_slot, parent_slot.data(), parent_slot.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess && !level_nodes.empty())
            e = hipMemcpy(d.level_nodes, level_nodes.data(), level_nodes.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d.level_offsets, offs.data(), offs.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(d.pair_counts, 0, ((size_t)n + 1) * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(d.pair_offsets, 0, ((size_t)n + 1) * sizeof(unsigned long long));
        if (e != hipSuccess) rc = hip_fail(e, "Bvh upload");
    }
    if (rc != PQ_OK) {
        pq_bvh_destroy(b);
        return rc;
    }
    *out = b;
    return PQ_OK;
}
}  // namespace

int pq_bvh_create(const float *aabbs, uint32_t n, pq_bvh **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_bvh_create: null output");
    *out = nullptr;
    if (n && !aabbs) return fail(PQ_ERR_INVALID, "pq_bvh_create: null aabbs");
    std::vector<LeafBox> leaves(n);
    if (n) std::memcpy(leaves.data(), aabbs, sizeof(LeafBox) * (size_t)n);
    std::vector<pq_bvh_node_wide> nodes;
    uint32_t depth = 0;
    build_binned_bvh(leaves, nodes, &depth);
    return bvh_handle_from_nodes(nodes, n, depth, out);
}

int pq_bvh_create_from_nodes(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint32_t num_leaves, pq_bvh **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_bvh_create_from_nodes: null output");
    *out = nullptr;
    if (num_nodes && !nodes) return fail(PQ_ERR_INVALID, "pq_bvh_create_from_nodes: null nodes");
    if (num_leaves == 0 && num_nodes == 0) return bvh_handle_from_nodes({}, 0, 0, out);
    bool ok = false;
    const uint32_t depth = bvh_depth(nodes, num_nodes, num_leaves, &ok);
    // every leaf 0..num_leaves-1 exactly once
    std::vector<uint8_t> seen(num_leaves, 0);
    uint64_t leaves = 0;
    for (uint32_t w = 0; ok && w < num_nodes; ++w)
        for (const pq_bvh_node *c : {&nodes[w].left, &nodes[w].right})
            if (is_leaf(*c)) {
                ok = c->children < num_leaves && !seen[c->children];
                if (ok) seen[c->children] = 1;
                ++leaves;
            }
    if (!ok || num_leaves == 0 || leaves != num_leaves)
        return fail(PQ_ERR_INVALID, "pq_bvh_create_from_nodes: not a well-formed depth-first tree over the leaves");
    return bvh_handle_from_nodes(std::vector<pq_bvh_node_wide>(nodes, nodes + num_nodes), num_leaves, depth, out);
}

size_t pq_bvh_ploc_workspace_bytes(uint32_t n) { return ploc_workspace_bytes(n); }

int pq_bvh_from_leaves_ploc(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t capacity,
                            uint32_t *num_nodes, uint32_t *depth, void *workspace, void *stream) {
    if (!num_nodes) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves_ploc: null num_nodes");
    *num_nodes = n == 0 ? 0u : (n == 1 ? 1u : n - 1);
    if (depth) *depth = 0;
    if (n == 0) return PQ_OK;
    if (!aabbs) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves_ploc: null aabbs");
    if (capacity < *num_nodes)
        return fail(PQ_ERR_CAPACITY, "pq_bvh_from_leaves_ploc: need " + std::to_string(*num_nodes) + " records");
    if (!out) return fail(PQ_ERR_INVALID, "pq_bvh_from_leaves_ploc: null output");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, ploc_workspace_bytes(n), &ws));
    uint32_t d = 0;
    PQ_HIP_CHECK(build_ploc_bvh(aabbs, n, out, &d, ws, (hipStream_t)stream));
    if (depth) *depth = d;
    return PQ_OK;
}

int pq_bvh_create_ploc(const float *aabbs, uint32_t n, pq_bvh **out, void *stream) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_bvh_create_ploc: null output");
    *out = nullptr;
    if (n && !aabbs) return fail(PQ_ERR_INVALID, "pq_bvh_create_ploc: null aabbs");
    const uint32_t nn = n == 0 ? 0u : (n == 1 ? 1u : n - 1);
    std::vector<pq_bvh_node_wide> nodes(nn);
    uint32_t depth = 0;
    if (n) {
        pq_bvh_node_wide *d_nodes = nullptr;
        hipError_t e = hipMalloc(&d_nodes, nn * sizeof(pq_bvh_node_wide));
        if (e != hipSuccess) return hip_fail(e, "pq_bvh_create_ploc");
        uint32_t got = 0;
        int rc = pq_bvh_from_leaves_ploc(aabbs, n, d_nodes, nn, &got, &depth, nullptr, stream);
        if (rc == PQ_OK) {
            e = hipMemcpy(nodes.data(), d_nodes, nn * sizeof(pq_bvh_node_wide), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(e, "pq_bvh_create_ploc download");
        }
        (void)hipFree(d_nodes);
        if (rc != PQ_OK) return rc;
    }
    return bvh_handle_from_nodes(nodes, n, depth, out);
}

int pq_bvh_destroy(pq_bvh *b) {
    if (!b) return PQ_OK;
    for (void *p : b->allocations) (void)hipFree(p);
    delete b;
    return PQ_OK;
}

int pq_bvh_info(const pq_bvh *b, uint32_t *num_leaves, uint32_t *num_nodes, uint32_t *depth) {
    if (!b) return fail(PQ_ERR_INVALID, "pq_bvh_info: null bvh");
    if (num_leaves) *num_leaves = b->dev.num_leaves;
    if (num_nodes) *num_nodes = b->dev.num_nodes;
    if (depth) *depth = b->dev.depth;
    return PQ_OK;
}

int pq_bvh_export(const pq_bvh *b, pq_bvh_node_wide *out, uint32_t capacity) {
    if (!b || !out) return fail(PQ_ERR_INVALID, "pq_bvh_export: null argument");
    if (capacity < b->dev.num_nodes)
        return fail(PQ_ERR_CAPACITY, "pq_bvh_export: need " + std::to_string(b->dev.num_nodes) + " records");
    PQ_HIP_CHECK(hipDeviceSynchronize());
    if (b->dev.num_nodes)
        PQ_HIP_CHECK(hipMemcpy(out, b->dev.nodes, b->dev.num_nodes * sizeof(pq_bvh_node_wide), hipMemcpyDeviceToHost));
    return PQ_OK;
}

int pq_bvh_update_and_refit(pq_bvh *b, const float *aabbs, void *stream) {
    if (!b) return fail(PQ_ERR_INVALID, "pq_bvh_update_and_refit: null bvh");
    if (b->dev.num_leaves && !aabbs) return fail(PQ_ERR_INVALID, "pq_bvh_update_and_refit: null aabbs");
    PQ_HIP_CHECK(launch_bvh_refit(b->dev, aabbs, (hipStream_t)stream));
    return PQ_OK;
}

int pq_bvh_self_pairs(pq_bvh *b, uint32_t *pairs, uint64_t capacity, uint64_t *count, void *stream) {
    if (!b || !count) return fail(PQ_ERR_INVALID, "pq_bvh_self_pairs: null argument");
    if (capacity && !pairs) return fail(PQ_ERR_INVALID, "pq_bvh_self_pairs: null pairs buffer");
    PQ_HIP_CHECK(launch_bvh_self_pairs(b->dev, pairs, capacity, (unsigned long long *)count, (hipStream_t)stream));
    return PQ_OK;
}

size_t pq_contact_workspace_bytes(uint64_t num_pairs) { return contact_workspace_bytes(num_pairs); }

int pq_contact_set_tuning(const pq_contact_tuning *t) {
    if (!t) return fail(PQ_ERR_INVALID, "pq_contact_set_tuning: null tuning");
    const uint32_t g = t->epa_lanes_per_pair;
    if ((g != 64 && g != 32 && g != 16 && g != 8 && g != 4) || t->epa_polytope > 4 || t->epa_schedule > 2 ||
        t->reserved)
        return fail(PQ_ERR_INVALID, "pq_contact_set_tuning: epa_lanes_per_pair must be 64, 32, 16, 8 or 4, "
                                    "epa_polytope 0 to 4 and epa_schedule 0 to 2");
    g_epa_lanes_per_pair = (int)g;
    g_epa_polytope = (int)t->epa_polytope;
    g_epa_schedule = (int)t->epa_schedule;
    return PQ_OK;
}

int pq_contact_set_phase_timing(int enable) {
    if (enable) {
        for (hipEvent_t &ev : g_contact_phase_events)
            if (!ev) PQ_HIP_CHECK(hipEventCreate(&ev));
    }
    g_contact_phase_timing = enable != 0;
    return PQ_OK;
}

int pq_contact_phase_ms(float *ms, uint32_t n) {
    if (!ms || n < (uint32_t)kContactPhases) return fail(PQ_ERR_INVALID, "pq_contact_phase_ms: need 5 floats");
    if (!g_contact_phase_events[kContactPhases])
        return fail(PQ_ERR_INVALID, "pq_contact_phase_ms: phase timing was never enabled");
    PQ_HIP_CHECK(hipEventSynchronize(g_contact_phase_events[kContactPhases]));
    for (int k = 0; k < kContactPhases; ++k)
        PQ_HIP_CHECK(hipEventElapsedTime(&ms[k], g_contact_phase_events[k], g_contact_phase_events[k + 1]));
    return PQ_OK;
}

int pq_contact_get_tuning(pq_contact_tuning *t) {
    if (!t) return fail(PQ_ERR_INVALID, "pq_contact_get_tuning: null tuning");
    *t = pq_contact_tuning{(uint32_t)g_epa_lanes_per_pair, (uint32_t)g_epa_polytope, (uint32_t)g_epa_schedule, 0};
    return PQ_OK;
}

int pq_contact_pairs(const pq_shapes *s, const pq_isometry *poses, const uint32_t *pairs, uint64_t m, float prediction,
                     const float *init_dirs, pq_contact *out, uint32_t *iters, void *workspace, void *stream) {
    if (!s) return fail(PQ_ERR_INVALID, "pq_contact_pairs: null shapes");
    if (m == 0) return PQ_OK;
    if (!poses || !pairs || !out) return fail(PQ_ERR_INVALID, "pq_contact_pairs: null buffer");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, contact_workspace_bytes(m), &ws));
    PQ_HIP_CHECK(launch_contacts(s->dev, (const float *)poses, pairs, m, nullptr, prediction, init_dirs, out, iters,
                                 ws, (hipStream_t)stream));
    return PQ_OK;
}

int pq_contact_pairs_device_count(const pq_shapes *s, const pq_isometry *poses, const uint32_t *pairs,
                                  uint64_t capacity, const uint64_t *num_pairs, float prediction,
                                  const float *init_dirs, pq_contact *out, uint32_t *iters, void *workspace,
                                  void *stream) {
    if (!s || !num_pairs) return fail(PQ_ERR_INVALID, "pq_contact_pairs_device_count: null argument");
    if (capacity == 0) return PQ_OK;
    if (!poses || !pairs || !out) return fail(PQ_ERR_INVALID, "pq_contact_pairs_device_count: null buffer");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, contact_workspace_bytes(capacity), &ws));
    PQ_HIP_CHECK(launch_contacts(s->dev, (const float *)poses, pairs, capacity, num_pairs, prediction, init_dirs, out,
                                 iters, ws, (hipStream_t)stream));
    return PQ_OK;
}

}  // extern "C"
