// contact_noalloc.hip -- pq_contact_pairs_noalloc: the batched query::contact of
// pq_contact_pairs with its workspace (EPA queue, tier lists, counters,
// continuation records, polytope staging) in one fixed device array, so the call
// makes no device allocation.  It is for callers that cannot allocate: code
// under hipGraph capture, serving loops, harnesses that forbid hipMalloc.
//
// The array lives in this file's code object, one copy per GPU, held until the
// process ends.  Its device address is looked up once per device (with that
// device current) and used for every launch; the array's name itself is only a
// host-side stand-in.  All calls on one GPU share that GPU's array: calls that
// overlap on one GPU (several streams, several host threads) race on it.  The
// contact launch clears its counters with a kernel on every call, and every
// other byte of the workspace is written before it is read, so nothing left
// over from an earlier call is seen.
#include <mutex>

#include "pq_internal.h"

namespace {

// 124.7 MiB of contact workspace covers 262,144 pairs.
constexpr uint64_t kNoAllocMaxPairs = PQ_CONTACT_NOALLOC_MAX_PAIRS;
constexpr size_t kScratchBytes = 128ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on each device.
int scratch_address(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pq::hip_fail(e, "pq_contact_pairs_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return pq::fail(PQ_ERR_INVALID, "pq_contact_pairs_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return pq::hip_fail(e, "pq_contact_pairs_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return pq::fail(PQ_ERR_HIP, "pq_contact_pairs_noalloc: hipGetSymbolAddress gave a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PQ_OK;
}

}  // namespace

extern "C" {

int pq_contact_pairs_noalloc(const pq_shapes *s, const pq_isometry *poses, const uint32_t *pairs, uint64_t m,
                             float prediction, const float *init_dirs, pq_contact *out, uint32_t *iters,
                             void *stream) {
    if (!s) return pq::fail(PQ_ERR_INVALID, "pq_contact_pairs_noalloc: null shapes");
    if (m > kNoAllocMaxPairs)
        return pq::fail(PQ_ERR_INVALID, "pq_contact_pairs_noalloc: " + std::to_string(m) +
                                            " pairs exceed the fixed workspace's maximum of " +
                                            std::to_string(kNoAllocMaxPairs) + " pairs");
    if (m == 0) return PQ_OK;
    if (!poses || !pairs || !out) return pq::fail(PQ_ERR_INVALID, "pq_contact_pairs_noalloc: null buffer");
    if (pq::contact_workspace_bytes(m) > kScratchBytes)
        return pq::fail(PQ_ERR_INVALID, "pq_contact_pairs_noalloc: workspace larger than the fixed array");
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != PQ_OK) return rc;
    PQ_HIP_CHECK(pq::launch_contacts(s->dev, (const float *)poses, pairs, m, nullptr, prediction, init_dirs, out,
                                     iters, ws, (hipStream_t)stream));
    return PQ_OK;
}

}  // extern "C"
