// dynamics.hip -- what a simulation frame adds around the collision queries:
// rigid-motion pose updates and a per-pair cache of the last contact normal,
// used as the next frame's GJK seed (as parry's manifold code reuses its
// previous normal, contact_manifolds_pfm_pfm.rs:67).  No counterpart in parry's
// query path.
#include "pq_internal.h"
#include "pq_math.h"

namespace pq {
namespace {

// Pose update with constant linear and angular velocity (world frame) over dt:
// t += v dt;  q = normalize(dq * q), dq the rotation by |w| dt about w / |w|.
__global__ void integrate_kernel(float *__restrict__ poses, const float *__restrict__ lin,
                                 const float *__restrict__ ang, uint32_t n, float dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *p = poses + 7 * (size_t)i;
    const f3 v{lin[3 * i], lin[3 * i + 1], lin[3 * i + 2]};
    const f3 w{ang[3 * i], ang[3 * i + 1], ang[3 * i + 2]};
    p[4] += v.x * dt;
    p[5] += v.y * dt;
    p[6] += v.z * dt;
    const float wn = sqrtf(norm_sq3(w));
    if (wn > 0.0f) {
        const float half = 0.5f * wn * dt;
        const float s = sinf(half) / wn, c = cosf(half);
        const quat dq{w.x * s, w.y * s, w.z * s, c};
        const quat q{p[0], p[1], p[2], p[3]};
        quat r = qmul(dq, q);
        const float nrm = sqrtf((r.x * r.x + r.y * r.y) + (r.z * r.z + r.w * r.w));
        p[0] = r.x / nrm;
        p[1] = r.y / nrm;
        p[2] = r.z / nrm;
        p[3] = r.w / nrm;
    }
}

// Open-addressing table of (i, j) -> unit normal in shape i's frame; key 0 = empty
// (keys are (i + 1) << 32 | j, never 0).
struct CacheSlot {
    unsigned long long key;
    float n[3];
    uint32_t pad;
};
static_assert(sizeof(CacheSlot) == 24, "cache slot");

__device__ __forceinline__ uint32_t slot_hash(unsigned long long key, uint32_t mask) {
    unsigned long long h = key * 0x9E3779B97F4A7C15ull;
    return (uint32_t)(h >> 32) & mask;
}

__global__ void cache_clear_kernel(CacheSlot *__restrict__ t, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) t[i].key = 0ull;
}

__device__ __forceinline__ uint64_t device_count(const unsigned long long *m_dev, uint64_t cap) {
    const uint64_t m = *m_dev;
    return m < cap ? m : cap;
}

__global__ void cache_store_kernel(CacheSlot *__restrict__ t, uint32_t cap, const uint32_t *__restrict__ pairs,
                                   const unsigned long long *__restrict__ m_dev, uint64_t pair_cap,
                                   const float *__restrict__ contacts, const float *__restrict__ poses) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= device_count(m_dev, pair_cap)) return;
    const float *c = contacts + 14 * k;
    if (__float_as_uint(c[13]) != 1u) return;
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const f3 n = qrot_inv(load_iso(poses + 7 * (size_t)i).q, f3{c[6], c[7], c[8]});
    const unsigned long long key = ((unsigned long long)(i + 1) << 32) | j;
    const uint32_t mask = cap - 1;
    uint32_t h = slot_hash(key, mask);
    for (uint32_t probe = 0; probe < cap; ++probe, h = (h + 1) & mask) {
        const unsigned long long prev = atomicCAS(&t[h].key, 0ull, key);
        if (prev == 0ull || prev == key) {  // pairs are unique: the slot is ours
            t[h].n[0] = n.x;
            t[h].n[1] = n.y;
            t[h].n[2] = n.z;
            return;
        }
    }
}

__global__ void cache_apply_kernel(const CacheSlot *__restrict__ t, uint32_t cap, const uint32_t *__restrict__ pairs,
                                   const unsigned long long *__restrict__ m_dev, uint64_t pair_cap,
                                   float *__restrict__ seeds, uint32_t *__restrict__ hits) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= device_count(m_dev, pair_cap)) return;
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const unsigned long long key = ((unsigned long long)(i + 1) << 32) | j;
    const uint32_t mask = cap - 1;
    uint32_t h = slot_hash(key, mask);
    for (uint32_t probe = 0; probe < cap; ++probe, h = (h + 1) & mask) {
        const unsigned long long kk = t[h].key;
        if (kk == 0ull) return;
        if (kk == key) {
            seeds[3 * k] = t[h].n[0];
            seeds[3 * k + 1] = t[h].n[1];
            seeds[3 * k + 2] = t[h].n[2];
            if (hits) atomicAdd(hits, 1u);
            return;
        }
    }
}

__global__ void zero_kernel(uint32_t *__restrict__ p, uint64_t words) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0u;
}

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace
}  // namespace pq

using namespace pq;

extern "C" {

int pq_integrate_poses(pq_isometry *poses, const float *linvel, const float *angvel, uint32_t n, float dt,
                       void *stream) {
    if (n && (!poses || !linvel || !angvel)) return fail(PQ_ERR_INVALID, "pq_integrate_poses: null argument");
    if (n == 0) return PQ_OK;
    hipLaunchKernelGGL(integrate_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, (float *)poses,
                       linvel, angvel, n, dt);
    PQ_HIP_CHECK(hipGetLastError());
    return PQ_OK;
}

int pq_device_zero(void *ptr, size_t bytes, void *stream) {
    if (bytes % 4 != 0 || ((uintptr_t)ptr & 3u) != 0) return fail(PQ_ERR_INVALID, "pq_device_zero: 4-byte units only");
    if (bytes == 0) return PQ_OK;
    if (!ptr) return fail(PQ_ERR_INVALID, "pq_device_zero: null pointer");
    hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(bytes / 4)), dim3(256), 0, (hipStream_t)stream, (uint32_t *)ptr,
                       (uint64_t)(bytes / 4));
    PQ_HIP_CHECK(hipGetLastError());
    return PQ_OK;
}

size_t pq_warm_cache_bytes(uint32_t capacity) { return (size_t)capacity * sizeof(CacheSlot); }

int pq_warm_cache_clear(void *cache, uint32_t capacity, void *stream) {
    if (!cache || capacity == 0 || (capacity & (capacity - 1)))
        return fail(PQ_ERR_INVALID, "pq_warm_cache_clear: capacity must be a power of two");
    hipLaunchKernelGGL(cache_clear_kernel, dim3(blocks_for(capacity)), dim3(256), 0, (hipStream_t)stream,
                       (CacheSlot *)cache, capacity);
    PQ_HIP_CHECK(hipGetLastError());
    return PQ_OK;
}

int pq_warm_cache_store(void *cache, uint32_t capacity, const uint32_t *pairs, const uint64_t *num_pairs,
                        uint64_t pair_capacity, const pq_contact *contacts, const pq_isometry *poses, void *stream) {
    if (!cache || !pairs || !num_pairs || !contacts || !poses || capacity == 0 || (capacity & (capacity - 1)))
        return fail(PQ_ERR_INVALID, "pq_warm_cache_store: invalid argument");
    if (pair_capacity == 0) return PQ_OK;
    hipLaunchKernelGGL(cache_store_kernel, dim3(blocks_for(pair_capacity)), dim3(256), 0, (hipStream_t)stream,
                       (CacheSlot *)cache, capacity, pairs, (const unsigned long long *)num_pairs, pair_capacity,
                       (const float *)contacts, (const float *)poses);
    PQ_HIP_CHECK(hipGetLastError());
    return PQ_OK;
}

int pq_warm_cache_apply(const void *cache, uint32_t capacity, const uint32_t *pairs, const uint64_t *num_pairs,
                        uint64_t pair_capacity, float *seeds, uint32_t *hits, void *stream) {
    if (!cache || !pairs || !num_pairs || !seeds || capacity == 0 || (capacity & (capacity - 1)))
        return fail(PQ_ERR_INVALID, "pq_warm_cache_apply: invalid argument");
    if (pair_capacity == 0) return PQ_OK;
    hipLaunchKernelGGL(cache_apply_kernel, dim3(blocks_for(pair_capacity)), dim3(256), 0, (hipStream_t)stream,
                       (const CacheSlot *)cache, capacity, pairs, (const unsigned long long *)num_pairs,
                       pair_capacity, seeds, hits);
    PQ_HIP_CHECK(hipGetLastError());
    return PQ_OK;
}

int pq_stream_begin_capture(void *stream) {
    if (!stream) return fail(PQ_ERR_INVALID, "pq_stream_begin_capture: the legacy default stream cannot be captured");
    PQ_HIP_CHECK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return PQ_OK;
}

int pq_stream_end_capture(void *stream, pq_graph **out) {
    if (!stream || !out) return fail(PQ_ERR_INVALID, "pq_stream_end_capture: null argument");
    *out = nullptr;
    auto *g = new pq_graph();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        pq_graph_destroy(g);
        return hip_fail(e, "pq_stream_end_capture");
    }
    *out = g;
    return PQ_OK;
}

}  // extern "C"
