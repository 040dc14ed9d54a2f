// capi_mlp.cpp -- C ABI of the learned GJK warm start (no counterpart in parry).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mlp_internal.h"

using namespace pq;
using namespace pq::mlp;

struct pq_mlp {
    DeviceMlp dev{};
    TrainWork work{};
    std::vector<void *> allocations;
    std::vector<void *> work_allocations;
    uint64_t adam_t = 0;  // Adam steps taken
};

namespace {

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// He initialisation: W ~ N(0, 2 / fan_in) (Box-Muller), b = 0.
std::vector<float> he_init(uint64_t seed) {
    std::vector<float> p(kNumParams, 0.0f);
    uint64_t s = seed;
    const uint32_t dims[4][2] = {{kH, kIn}, {kH, kH}, {kH, kH}, {kOut, kH}};
    for (uint32_t l = 0; l < 4; ++l) {
        const double sd = std::sqrt(2.0 / dims[l][1]);
        for (uint32_t k = 0; k < dims[l][0] * dims[l][1]; ++k) {
            const double u1 = std::max(uniform01(s), 1e-300), u2 = uniform01(s);
            p[w_offset(l) + k] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    return p;
}

int net_alloc(std::vector<void *> &list, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    list.push_back(*p);
    return PQ_OK;
}

int upload_params(pq_mlp *net, const float *host) {
    PQ_HIP_CHECK(hipMemcpy(net->dev.params, host, kNumParams * sizeof(float), hipMemcpyHostToDevice));
    PQ_HIP_CHECK(hipMemset(net->dev.m1, 0, kNumParams * sizeof(float)));
    PQ_HIP_CHECK(hipMemset(net->dev.m2, 0, kNumParams * sizeof(float)));
    net->adam_t = 0;
    PQ_HIP_CHECK(launch_pack(net->dev, nullptr));
    PQ_HIP_CHECK(hipDeviceSynchronize());
    return PQ_OK;
}

// Training buffers for batches of up to `batch` pairs (grown on demand; zeroed,
// so the padding rows of the output delta stay zero).
int ensure_work(pq_mlp *net, uint32_t batch) {
    const uint32_t need = (batch + 31) / 32 * 32;
    if (net->work.max_batch >= need) return PQ_OK;
    PQ_HIP_CHECK(hipDeviceSynchronize());
    for (void *p : net->work_allocations) (void)hipFree(p);
    net->work_allocations.clear();
    net->work = TrainWork{};
    TrainWork w{};
    w.max_batch = need;
    const size_t row = (size_t)need * sizeof(uint16_t);
    uint16_t **bufs[8] = {&w.xb, &w.h1, &w.h2, &w.h3, &w.d4, &w.d3, &w.d2, &w.d1};
    const uint32_t rows[8] = {64, kH, kH, kH, 16, kH, kH, kH};
    for (int k = 0; k < 8; ++k) {
        int rc = net_alloc(net->work_allocations, (void **)bufs[k], rows[k] * row);
        if (rc != PQ_OK) return rc;
        PQ_HIP_CHECK(hipMemset(*bufs[k], 0, rows[k] * row));
    }
    int rc = net_alloc(net->work_allocations, (void **)&w.loss, need * sizeof(float));
    if (rc == PQ_OK) {
        const size_t chunks = (need + kChunk - 1) / kChunk;
        rc = net_alloc(net->work_allocations, (void **)&w.partial, chunks * kNumParams * sizeof(float));
    }
    if (rc == PQ_OK) rc = net_alloc(net->work_allocations, (void **)&w.loss_total, sizeof(double));
    if (rc != PQ_OK) return rc;
    PQ_HIP_CHECK(hipMemset(w.loss_total, 0, sizeof(double)));
    net->work = w;
    return PQ_OK;
}

}  // namespace

extern "C" {

int pq_mlp_create(const float *params, uint64_t seed, pq_mlp **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_mlp_create: null output");
    *out = nullptr;
    auto *net = new pq_mlp();
    DeviceMlp &d = net->dev;
    int rc = net_alloc(net->allocations, (void **)&d.params, kNumParams * sizeof(float));
    if (rc == PQ_OK) rc = net_alloc(net->allocations, (void **)&d.m1, kNumParams * sizeof(float));
    if (rc == PQ_OK) rc = net_alloc(net->allocations, (void **)&d.m2, kNumParams * sizeof(float));
    if (rc == PQ_OK) rc = net_alloc(net->allocations, (void **)&d.fw, kFwdFrags * 64 * sizeof(uint4));
    if (rc == PQ_OK) rc = net_alloc(net->allocations, (void **)&d.bw, kBwdFrags * 64 * sizeof(uint4));
    if (rc == PQ_OK) rc = net_alloc(net->allocations, (void **)&d.bias, kBiasCount * sizeof(float));
    if (rc == PQ_OK) {
        if (params) {
            rc = upload_params(net, params);
        } else {
            const std::vector<float> p = he_init(seed);
            rc = upload_params(net, p.data());
        }
    }
    if (rc != PQ_OK) {
        pq_mlp_destroy(net);
        return rc;
    }
    *out = net;
    return PQ_OK;
}

int pq_mlp_destroy(pq_mlp *net) {
    if (!net) return PQ_OK;
    (void)hipDeviceSynchronize();
    for (void *p : net->work_allocations) (void)hipFree(p);
    for (void *p : net->allocations) (void)hipFree(p);
    delete net;
    return PQ_OK;
}

int pq_mlp_get_params(const pq_mlp *net, float *params) {
    if (!net || !params) return fail(PQ_ERR_INVALID, "pq_mlp_get_params: null argument");
    PQ_HIP_CHECK(hipDeviceSynchronize());
    PQ_HIP_CHECK(hipMemcpy(params, net->dev.params, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return PQ_OK;
}

int pq_mlp_set_params(pq_mlp *net, const float *params) {
    if (!net || !params) return fail(PQ_ERR_INVALID, "pq_mlp_set_params: null argument");
    PQ_HIP_CHECK(hipDeviceSynchronize());
    return upload_params(net, params);
}

int pq_shape_descriptors(const pq_shapes *shapes, float *out, void *stream) {
    if (!shapes || (!out && shapes->dev.n)) return fail(PQ_ERR_INVALID, "pq_shape_descriptors: null argument");
    PQ_HIP_CHECK(launch_descriptors(shapes->dev, out, (hipStream_t)stream));
    return PQ_OK;
}

int pq_pair_features(const pq_isometry *poses, const uint32_t *pairs, uint64_t num_pairs, const float *descriptors,
                     float *features, void *stream) {
    if (num_pairs && (!poses || !pairs || !descriptors || !features))
        return fail(PQ_ERR_INVALID, "pq_pair_features: null argument");
    PQ_HIP_CHECK(launch_features((const float *)poses, pairs, num_pairs, descriptors, features, (hipStream_t)stream));
    return PQ_OK;
}

int pq_mlp_forward(const pq_mlp *net, const float *features, uint64_t n, float *out, void *stream) {
    if (!net || (n && (!features || !out))) return fail(PQ_ERR_INVALID, "pq_mlp_forward: null argument");
    PQ_HIP_CHECK(launch_forward_rows(net->dev, features, n, out, (hipStream_t)stream));
    return PQ_OK;
}

int pq_contact_seeds(const pq_mlp *net, const pq_isometry *poses, const uint32_t *pairs, uint64_t num_pairs,
                     const float *descriptors, float *seeds, void *stream) {
    if (!net || (num_pairs && (!poses || !pairs || !descriptors || !seeds)))
        return fail(PQ_ERR_INVALID, "pq_contact_seeds: null argument");
    PQ_HIP_CHECK(launch_seeds(net->dev, (const float *)poses, pairs, num_pairs, nullptr, descriptors, seeds,
                              (hipStream_t)stream));
    return PQ_OK;
}

int pq_contact_seeds_device_count(const pq_mlp *net, const pq_isometry *poses, const uint32_t *pairs,
                                  uint64_t capacity, const uint64_t *num_pairs, const float *descriptors, float *seeds,
                                  void *stream) {
    if (!net || !num_pairs || (capacity && (!poses || !pairs || !descriptors || !seeds)))
        return fail(PQ_ERR_INVALID, "pq_contact_seeds_device_count: null argument");
    PQ_HIP_CHECK(launch_seeds(net->dev, (const float *)poses, pairs, capacity, num_pairs, descriptors, seeds,
                              (hipStream_t)stream));
    return PQ_OK;
}

int pq_mlp_gradients(pq_mlp *net, const float *features, const float *labels, const uint32_t *rows, uint32_t count,
                     float *grads, double *loss_sum, void *stream) {
    if (!net || !features || !labels || !rows || !grads || count == 0)
        return fail(PQ_ERR_INVALID, "pq_mlp_gradients: null argument or empty batch");
    if (count > 65536) return fail(PQ_ERR_INVALID, "pq_mlp_gradients: at most 65536 rows per batch");
    int rc = ensure_work(net, count);
    if (rc != PQ_OK) return rc;
    const hipStream_t s = (hipStream_t)stream;
    PQ_HIP_CHECK(hipMemsetAsync(net->work.loss_total, 0, sizeof(double), s));
    PQ_HIP_CHECK(launch_train_step(net->dev, net->work, features, labels, rows, count, nullptr, grads, s));
    if (loss_sum)
        PQ_HIP_CHECK(hipMemcpyAsync(loss_sum, net->work.loss_total, sizeof(double), hipMemcpyDeviceToDevice, s));
    return PQ_OK;
}

int pq_mlp_train(pq_mlp *net, const float *features, const float *labels, uint64_t n, const pq_mlp_train_config *cfg,
                 double *epoch_losses, void *stream) {
    if (!net || !features || !labels || !cfg || n == 0) return fail(PQ_ERR_INVALID, "pq_mlp_train: null argument");
    if (cfg->batch == 0 || cfg->batch % 32 != 0 || cfg->batch > 65536 || cfg->reserved != 0 || n > 0xffffffffull)
        return fail(PQ_ERR_INVALID, "pq_mlp_train: batch must be a multiple of 32 up to 65536");
    int rc = ensure_work(net, cfg->batch);
    if (rc != PQ_OK) return rc;
    const hipStream_t s = (hipStream_t)stream;
    uint32_t *d_perm = nullptr;
    PQ_HIP_CHECK(hipMalloc((void **)&d_perm, n * sizeof(uint32_t)));
    std::vector<uint32_t> perm(n);
    uint64_t rng = cfg->shuffle_seed;
    const uint64_t steps_per_epoch = (n + cfg->batch - 1) / cfg->batch;
    const uint64_t total_steps = steps_per_epoch * cfg->epochs;
    uint64_t step = 0;
    hipError_t e = hipSuccess;
    for (uint32_t ep = 0; ep < cfg->epochs && e == hipSuccess; ++ep) {
        for (uint64_t k = 0; k < n; ++k) perm[k] = (uint32_t)k;
        for (uint64_t k = n - 1; k > 0; --k) std::swap(perm[k], perm[splitmix64(rng) % (k + 1)]);
        e = hipStreamSynchronize(s);  // the previous epoch's reads of d_perm are done
        if (e == hipSuccess) e = hipMemcpy(d_perm, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(net->work.loss_total, 0, sizeof(double), s);
        for (uint64_t b = 0; b < n && e == hipSuccess; b += cfg->batch, ++step) {
            const uint32_t count = (uint32_t)std::min<uint64_t>(cfg->batch, n - b);
            // cosine decay from lr to lr_final over all steps
            const double frac = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 0.0;
            const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(3.141592653589793 * frac));
            ++net->adam_t;
            const AdamStep st{(float)lr, cfg->beta1, cfg->beta2, cfg->eps,
                              (float)(1.0 - std::pow((double)cfg->beta1, (double)net->adam_t)),
                              (float)(1.0 - std::pow((double)cfg->beta2, (double)net->adam_t))};
            e = launch_train_step(net->dev, net->work, features, labels, d_perm + b, count, &st, nullptr, s);
        }
        double total = 0.0;
        if (e == hipSuccess) e = hipMemcpyAsync(&total, net->work.loss_total, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess && epoch_losses) epoch_losses[ep] = total / (double)n;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_perm);
    if (e != hipSuccess) return hip_fail(e, "pq_mlp_train");
    return PQ_OK;
}

int pq_contact_local_normals(const pq_contact *contacts, const pq_isometry *poses, const uint32_t *pairs,
                             uint64_t num_pairs, float *normals, void *stream) {
    if (num_pairs && (!contacts || !poses || !pairs || !normals))
        return fail(PQ_ERR_INVALID, "pq_contact_local_normals: null argument");
    PQ_HIP_CHECK(launch_local_normals((const float *)contacts, (const float *)poses, pairs, num_pairs, normals,
                                      (hipStream_t)stream));
    return PQ_OK;
}

}  // extern "C"
