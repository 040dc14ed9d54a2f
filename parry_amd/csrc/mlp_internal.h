// mlp_internal.h -- the warm-start MLP's device layout, shared by mlp.hip and capi_mlp.cpp.
#pragma once

#include "pq_internal.h"

namespace pq {
namespace mlp {

constexpr uint32_t kIn = 44;           // 12 relative-pose + 2 x 16 descriptor floats
constexpr uint32_t kFeatStride = 48;   // feature row stride (floats 44..47 are zero)
constexpr uint32_t kDesc = 16;         // floats per shape descriptor
constexpr uint32_t kH = 128;           // hidden width
constexpr uint32_t kOut = 3;
constexpr uint32_t kTiles = kH / 16;   // 16-unit output tiles of a hidden layer
constexpr uint32_t kNumParams = 128 * 44 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3;
static_assert(kNumParams == 39171, "parameter count");

// Offsets of W_l (row-major out x in) and b_l in the flat parameter vector.
__host__ __device__ constexpr uint32_t w_offset(uint32_t layer) {
    return layer == 0 ? 0u : layer == 1 ? 5760u : layer == 2 ? 22272u : 38784u;
}
__host__ __device__ constexpr uint32_t b_offset(uint32_t layer) {
    return layer == 0 ? 5632u : layer == 1 ? 22144u : layer == 2 ? 38656u : 39168u;
}

// Forward fragments (one uint4 = 8 bf16 per lane, 64 lanes per fragment):
// layer 0: 8 m-tiles x 2 k-blocks, layers 1-2: 8 x 4, layer 3: 1 x 4.
constexpr uint32_t kFragL0 = 0, kFragL1 = 16, kFragL2 = 48, kFragL3 = 80, kFwdFrags = 84;
// Transposed fragments for the backward pass: W4^T (8 x 1), W3^T (8 x 4), W2^T (8 x 4).
constexpr uint32_t kBwdL3 = 0, kBwdL2 = 8, kBwdL1 = 40, kBwdFrags = 72;
constexpr uint32_t kBias0 = 0, kBias1 = 128, kBias2 = 256, kBias3 = 384, kBiasCount = 387;
constexpr uint32_t kInferBlock = 512;  // threads per inference block (weights in LDS)
constexpr uint32_t kChunk = 1024;      // pairs per weight-gradient task
constexpr uint32_t kTileTasks = 32 + 64 + 64 + 8;
// |out|^2 below which a predicted seed is rejected (the contact kernel then
// uses parry's default seed).
constexpr float kSeedMinSq = 1.4210855e-14f;  // f32::EPSILON^2

// K position of slot s of lane group g inside a 32-unit block built from two
// 16-unit output tiles: slots 0-3 from the first tile, 4-7 from the second.
__host__ __device__ constexpr uint32_t kloc(uint32_t g, uint32_t s) { return s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4); }

struct DeviceMlp {
    float *params;  // f32 master parameters (kNumParams)
    float *m1, *m2; // Adam moments
    uint4 *fw;      // kFwdFrags * 64
    uint4 *bw;      // kBwdFrags * 64
    float *bias;    // kBiasCount
};

struct AdamStep {
    float lr, beta1, beta2, eps, bc1, bc2;
};

// Per-batch training buffers ([unit][pair] bf16 rows of stride max_batch).
struct TrainWork {
    uint16_t *xb, *h1, *h2, *h3, *d4, *d3, *d2, *d1;
    float *loss;
    float *partial;       // ceil(max_batch / kChunk) * kNumParams
    double *loss_total;   // running loss sum
    uint32_t max_batch;   // multiple of 32
};

hipError_t launch_descriptors(const DeviceShapes &s, float *out, hipStream_t stream);
hipError_t launch_features(const float *poses, const uint32_t *pairs, uint64_t m, const float *desc, float *out,
                           hipStream_t stream);
// m_dev (nullable, device u64): the pair count, capped by m.
hipError_t launch_seeds(const DeviceMlp &net, const float *poses, const uint32_t *pairs, uint64_t m,
                        const uint64_t *m_dev, const float *desc, float *seeds, hipStream_t stream);
hipError_t launch_forward_rows(const DeviceMlp &net, const float *X, uint64_t n, float *out, hipStream_t stream);
hipError_t launch_pack(const DeviceMlp &net, hipStream_t stream);
// One batch: forward, loss, backward; Adam step (adam != null) or gradient into grad_out.
hipError_t launch_train_step(const DeviceMlp &net, const TrainWork &w, const float *X, const float *Y,
                             const uint32_t *idx, uint32_t count, const AdamStep *adam, float *grad_out,
                             hipStream_t stream);
hipError_t launch_local_normals(const float *contacts, const float *poses, const uint32_t *pairs, uint64_t m,
                                float *out, hipStream_t stream);

}  // namespace mlp
}  // namespace pq
