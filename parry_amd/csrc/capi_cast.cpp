// capi_cast.cpp -- C ABI of the linear shape casts and of the contact manifolds.
#include <cstdint>

#include "pq_internal.h"

using namespace pq;

extern "C" {

size_t pq_cast_shapes_workspace_bytes(uint64_t num_pairs) { return cast_workspace_bytes(num_pairs); }

int pq_cast_shapes_pairs(const pq_shapes *s, const pq_isometry *poses, const float *vels, const uint32_t *pairs,
                         uint64_t m, const pq_shape_cast_options *options, pq_shape_cast_hit *out, uint32_t *iters,
                         void *workspace, void *stream) {
    if (!s || !options) return fail(PQ_ERR_INVALID, "pq_cast_shapes_pairs: null shapes or options");
    if (m == 0) return PQ_OK;
    if (!poses || !vels || !pairs || !out) return fail(PQ_ERR_INVALID, "pq_cast_shapes_pairs: null buffer");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, cast_workspace_bytes(m), &ws));
    PQ_HIP_CHECK(launch_casts(s->dev, (const float *)poses, vels, pairs, m, *options, out, iters, ws,
                              (hipStream_t)stream));
    return PQ_OK;
}

size_t pq_contact_manifolds_workspace_bytes(uint64_t num_pairs) { return manifold_workspace_bytes(num_pairs); }

int pq_contact_manifolds_pfm_pairs(const pq_shapes *s, const pq_isometry *poses, const uint32_t *pairs, uint64_t m,
                                   float prediction, pq_contact_manifold *out, void *workspace, void *stream) {
    if (!s) return fail(PQ_ERR_INVALID, "pq_contact_manifolds_pfm_pairs: null shapes");
    if (m == 0) return PQ_OK;
    if (!poses || !pairs || !out) return fail(PQ_ERR_INVALID, "pq_contact_manifolds_pfm_pairs: null buffer");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, manifold_workspace_bytes(m), &ws));
    PQ_HIP_CHECK(launch_manifolds(s->dev, (const float *)poses, pairs, m, prediction, out, ws, (hipStream_t)stream));
    return PQ_OK;
}

int pq_contact_manifolds_pfm_update(const pq_shapes *s, const pq_isometry *poses, const uint32_t *pairs, uint64_t m,
                                    float prediction, pq_contact_manifold *manifolds, void *workspace, void *stream) {
    if (!s) return fail(PQ_ERR_INVALID, "pq_contact_manifolds_pfm_update: null shapes");
    if (m == 0) return PQ_OK;
    if (!poses || !pairs || !manifolds) return fail(PQ_ERR_INVALID, "pq_contact_manifolds_pfm_update: null buffer");
    void *ws = workspace;
    if (!ws) PQ_HIP_CHECK(stream_workspace((hipStream_t)stream, manifold_workspace_bytes(m), &ws));
    PQ_HIP_CHECK(launch_manifolds_update(s->dev, (const float *)poses, pairs, m, prediction, manifolds, ws,
                                         (hipStream_t)stream));
    return PQ_OK;
}

}  // extern "C"
