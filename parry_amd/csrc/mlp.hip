// mlp.hip -- the learned GJK warm start on gfx950: a 44-128-128-128-3 MLP whose
// output direction seeds GJK (the init_dir slot of
// contact_support_map_support_map_with_params, contact_support_map_support_map.rs:52-60).
// No counterpart in parry.
//
// All matrix products run on v_mfma_f32_16x16x32_bf16 (bf16 operands, f32
// accumulation) in the transposed form H_out[unit][pair] = W[unit][in] * H_in[in][pair]:
// the weights are the A operand (M = output units), a wave's 16-pair tile is the
// N dimension.  The 16x16 f32 result tile leaves lane (g = lane/16, p = lane%16)
// holding units 4g..4g+3 of pair p, which is exactly what that lane must supply
// as B operand for the next layer when a 32-unit K block is formed from two
// adjacent output tiles (slots 0-3 from the first, 4-7 from the second).  The
// weights are packed on the device into that k order, so activations never
// leave registers between layers.
//
// Lane layout of the instruction (checked on the device): A[m = l%16][k = 8(l/16)+s],
// B[k = 8(l/16)+s][n = l%16], D[m = 4(l/16)+i][n = l%16].
#include <cstring>

#include "mlp_internal.h"
#include "pq_math.h"

namespace pq {
namespace mlp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// float -> bfloat16 bits, round to nearest even (NaN -> quiet NaN).
__host__ __device__ __forceinline__ uint32_t f2bf(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ uint32_t pack2(float lo, float hi) { return f2bf(lo) | (f2bf(hi) << 16); }
__device__ __forceinline__ uint4 pack8(const float *v) {
    return uint4{pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7])};
}
__device__ __forceinline__ float unpack_slot(const uint4 &q, int s) {
    const uint32_t w = s < 2 ? q.x : (s < 4 ? q.y : (s < 6 ? q.z : q.w));
    return bf2f((s & 1) ? (w >> 16) : (w & 0xffffu));
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// ---- the forward pass of one wave over 32 pairs (two 16-pair N tiles) ---------
// x[kb][nt]: layer-1 B fragments (features 32kb + 8g + s of pair p).
// hb1..hb3[kb][nt]: the hidden activations as B fragments of the next layer
// (kept for the backward pass).  out[nt]: lanes g = 0 hold the 3 outputs (slots 0-2).
struct Acts {
    uint4 hb1[4][2], hb2[4][2], hb3[4][2];
};

template <int KB, class FragPtr>
__device__ __forceinline__ void hidden_layer(const FragPtr frags, int frag_base, const float *bias,
                                             const uint4 (*in)[2], uint4 (*out)[2]) {
    constexpr int kblocks = KB;
    const uint32_t l = lane_id(), g = l >> 4;
    f32x4 acc[kTiles][2];
#pragma unroll
    for (int tm = 0; tm < kTiles; ++tm) acc[tm][0] = acc[tm][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < kblocks; ++kb) {
#pragma unroll
        for (int tm = 0; tm < kTiles; ++tm) {
            const uint4 a = frags[(frag_base + tm * kblocks + kb) * 64 + l];
            acc[tm][0] = mfma(a, in[kb][0], acc[tm][0]);
            acc[tm][1] = mfma(a, in[kb][1], acc[tm][1]);
        }
    }
#pragma unroll
    for (int tm = 0; tm < kTiles; ++tm) {
        float h[2][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float b = bias[16 * tm + 4 * g + i];
            h[0][i] = fmaxf(acc[tm][0][i] + b, 0.0f);
            h[1][i] = fmaxf(acc[tm][1][i] + b, 0.0f);
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const uint32_t lo = pack2(h[nt][0], h[nt][1]), hi = pack2(h[nt][2], h[nt][3]);
            if (tm & 1) {
                out[tm >> 1][nt].z = lo;
                out[tm >> 1][nt].w = hi;
            } else {
                out[tm >> 1][nt].x = lo;
                out[tm >> 1][nt].y = hi;
            }
        }
    }
}

template <class FragPtr>
__device__ __forceinline__ void forward_wave(const FragPtr frags, const float *bias, const uint4 (*x)[2], Acts &a,
                                             f32x4 out[2]) {
    hidden_layer<2>(frags, kFragL0, bias + kBias0, x, a.hb1);
    hidden_layer<4>(frags, kFragL1, bias + kBias1, a.hb1, a.hb2);
    hidden_layer<4>(frags, kFragL2, bias + kBias2, a.hb2, a.hb3);
    const uint32_t l = lane_id(), g = l >> 4;
    out[0] = out[1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const uint4 w = frags[(kFragL3 + kb) * 64 + l];
        out[0] = mfma(w, a.hb3[kb][0], out[0]);
        out[1] = mfma(w, a.hb3[kb][1], out[1]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t u = 4 * g + i;
        const float b = u < 3 ? bias[kBias3 + u] : 0.0f;
        out[0][i] += b;
        out[1][i] += b;
    }
}

// ---- features ---------------------------------------------------------------
// The 44 inputs of a pair (i, j): the rotation matrix (row-major) and the
// translation of pos12 = pos_i^-1 * pos_j, then the 16-float descriptors of
// shape i and shape j.  Lane (g, p) supplies features 8g..8g+7 (K block 0) and
// 32+8g..32+8g+7 (K block 1; features 44..63 are zero).
__device__ __forceinline__ void rot_matrix(const quat &q, float R[9]) {
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    R[0] = 1.0f - 2.0f * (y * y + z * z);
    R[1] = 2.0f * (x * y - z * w);
    R[2] = 2.0f * (x * z + y * w);
    R[3] = 2.0f * (x * y + z * w);
    R[4] = 1.0f - 2.0f * (x * x + z * z);
    R[5] = 2.0f * (y * z - x * w);
    R[6] = 2.0f * (x * z - y * w);
    R[7] = 2.0f * (y * z + x * w);
    R[8] = 1.0f - 2.0f * (x * x + y * y);
}

__device__ __forceinline__ void lane_features(const iso &pos12, const float *d1, const float *d2, uint32_t g,
                                              float x0[8], float x1[8]) {
    float R[9];
    rot_matrix(pos12.q, R);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        x1[s] = 0.0f;
        if (g == 0) {
            x0[s] = R[s];
            x1[s] = d2[4 + s];
        } else if (g == 1) {
            x0[s] = s == 0 ? R[8] : (s == 1 ? pos12.t.x : (s == 2 ? pos12.t.y : (s == 3 ? pos12.t.z : d1[s - 4])));
            if (s < 4) x1[s] = d2[12 + s];
        } else if (g == 2) {
            x0[s] = d1[4 + s];
        } else {
            x0[s] = s < 4 ? d1[12 + s] : d2[s - 4];
        }
    }
}

// ---- kernels -------------------------------------------------------------------

// Shape descriptors: [type one-hot (ball, cuboid, capsule, cone, polyhedron),
// params p0..p2, local AABB half extents (3), local AABB centre (3), bounding
// radius, vertex count / 64].
__global__ void descriptors_kernel(const ShapeRecord *__restrict__ rec, const float4 *__restrict__ verts, uint32_t n,
                                   float *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ShapeRecord r = rec[i];
    float d[kDesc];
    for (int k = 0; k < kDesc; ++k) d[k] = 0.0f;
    const float p0 = r.param[0], p1 = r.param[1], p2 = r.param[2];
    f3 he{0, 0, 0}, ce{0, 0, 0};
    float rad = 0.0f;
    switch (r.type) {
        case PQ_SHAPE_BALL:
            d[0] = 1.0f;
            he = f3{p0, p0, p0};
            rad = p0;
            break;
        case PQ_SHAPE_CUBOID:
            d[1] = 1.0f;
            he = f3{p0, p1, p2};
            rad = sqrtf(p0 * p0 + p1 * p1 + p2 * p2);
            break;
        case PQ_SHAPE_CAPSULE:
            d[2] = 1.0f;
            he = f3{p1, p0 + p1, p1};
            rad = p0 + p1;
            break;
        case PQ_SHAPE_CONE:
            d[3] = 1.0f;
            he = f3{p1, p0, p1};
            rad = sqrtf(p0 * p0 + p1 * p1);
            break;
        default: {
            d[4] = 1.0f;
            const float4 *v = verts + r.first_vertex;
            f3 mn{v[0].x, v[0].y, v[0].z}, mx = mn;
            float r2 = 0.0f;
            for (uint32_t k = 0; k < r.num_vertices; ++k) {
                const f3 q{v[k].x, v[k].y, v[k].z};
                mn = min3(mn, q);
                mx = max3(mx, q);
                r2 = fmaxf(r2, norm_sq3(q));
            }
            he = mul3(sub3(mx, mn), 0.5f);
            ce = mul3(add3(mx, mn), 0.5f);
            rad = sqrtf(r2);
            d[15] = (float)r.num_vertices / 64.0f;
        }
    }
    if (r.type != PQ_SHAPE_CONVEX_POLYHEDRON) {
        d[5] = p0;
        d[6] = p1;
        d[7] = p2;
    }
    d[8] = he.x; d[9] = he.y; d[10] = he.z;
    d[11] = ce.x; d[12] = ce.y; d[13] = ce.z;
    d[14] = rad;
    float4 *o = reinterpret_cast<float4 *>(out + (size_t)kDesc * i);
    for (int k = 0; k < 4; ++k) o[k] = float4{d[4 * k], d[4 * k + 1], d[4 * k + 2], d[4 * k + 3]};
}

// Feature rows (stride kFeatStride floats, zero padded) for training sets.
__global__ void features_kernel(const float *__restrict__ poses, const uint32_t *__restrict__ pairs, uint64_t m,
                                const float *__restrict__ desc, float *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const iso pos12 = iso_inv_mul(load_iso(poses + 7 * (size_t)i), load_iso(poses + 7 * (size_t)j));
    float R[9];
    rot_matrix(pos12.q, R);
    float *o = out + (size_t)kFeatStride * k;
    for (int s = 0; s < 9; ++s) o[s] = R[s];
    o[9] = pos12.t.x;
    o[10] = pos12.t.y;
    o[11] = pos12.t.z;
    for (int s = 0; s < kDesc; ++s) {
        o[12 + s] = desc[(size_t)kDesc * i + s];
        o[28 + s] = desc[(size_t)kDesc * j + s];
    }
    for (int s = kIn; s < kFeatStride; ++s) o[s] = 0.0f;
}

// Copies the forward fragments and the biases into LDS (block-cooperative).
__device__ __forceinline__ void stage_weights(const uint4 *__restrict__ fw, const float *__restrict__ bias,
                                              uint4 *lds_frag, float *lds_bias) {
    for (uint32_t k = threadIdx.x; k < kFwdFrags * 64; k += blockDim.x) lds_frag[k] = fw[k];
    for (uint32_t k = threadIdx.x; k < kBiasCount; k += blockDim.x) lds_bias[k] = bias[k];
    __syncthreads();
}

// Writes the output of pair p of tile nt (lanes g = 0): raw or normalised.
__device__ __forceinline__ void write_out(const f32x4 &o, uint64_t pair, uint64_t m, bool normalise, float *out) {
    if (pair >= m) return;
    float v0 = o[0], v1 = o[1], v2 = o[2];
    if (normalise) {
        const float sq = (v0 * v0 + v1 * v1) + v2 * v2;
        if (sq > kSeedMinSq && isfinite(sq)) {
            const float n = sqrtf(sq);
            v0 /= n;
            v1 /= n;
            v2 /= n;
        } else {
            v0 = v1 = v2 = 0.0f;  // the contact kernel then uses parry's default seed
        }
    }
    out[3 * pair] = v0;
    out[3 * pair + 1] = v1;
    out[3 * pair + 2] = v2;
}

// Fused inference over shape pairs: features, the four layers, normalised seeds.
// Persistent blocks of kInferBlock threads; each wave takes 32 pairs at a time.
__global__ __launch_bounds__(kInferBlock) void seeds_kernel(const uint4 *__restrict__ fw,
                                                            const float *__restrict__ bias,
                                                            const float *__restrict__ poses,
                                                            const uint32_t *__restrict__ pairs, uint64_t m,
                                                            const unsigned long long *__restrict__ m_dev,
                                                            const float *__restrict__ desc, float *__restrict__ seeds) {
    __shared__ uint4 lds_frag[kFwdFrags * 64];
    __shared__ float lds_bias[kBiasCount];
    if (m_dev) {  // pair count produced on the device (e.g. by the broadphase), capped by m
        const uint64_t md = *m_dev;
        if (md < m) m = md;
    }
    stage_weights(fw, bias, lds_frag, lds_bias);
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
    const uint64_t waves = (uint64_t)gridDim.x * (kInferBlock / 64);
    const uint64_t wave = (uint64_t)blockIdx.x * (kInferBlock / 64) + threadIdx.x / 64;
    for (uint64_t base = wave * 32; base < m; base += waves * 32) {
        // keeps the weight reads inside the loop (hoisted, they would not fit in registers)
        asm volatile("" ::: "memory");
        uint4 x[2][2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const uint64_t k = base + 16 * nt + p;
            float x0[8], x1[8];
            if (k < m) {
                const uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
                const iso pos12 = iso_inv_mul(load_iso(poses + 7 * (size_t)i), load_iso(poses + 7 * (size_t)j));
                lane_features(pos12, desc + (size_t)kDesc * i, desc + (size_t)kDesc * j, g, x0, x1);
            } else {
                for (int s = 0; s < 8; ++s) x0[s] = x1[s] = 0.0f;
            }
            x[0][nt] = pack8(x0);
            x[1][nt] = pack8(x1);
        }
        Acts a;
        f32x4 out[2];
        forward_wave(lds_frag, lds_bias, x, a, out);
        if (g == 0) {
            write_out(out[0], base + p, m, true, seeds);
            write_out(out[1], base + 16 + p, m, true, seeds);
        }
    }
}

// Loads the layer-1 B fragments of rows `idx` (or row = pair when idx is null).
__device__ __forceinline__ void load_rows(const float *__restrict__ X, const uint32_t *__restrict__ idx,
                                          uint64_t base, uint64_t count, uint4 x[2][2]) {
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const uint64_t k = base + 16 * nt + p;
        float x0[8], x1[8];
        for (int s = 0; s < 8; ++s) x0[s] = x1[s] = 0.0f;
        if (k < count) {
            const uint64_t row = idx ? idx[k] : k;
            const float4 *r = reinterpret_cast<const float4 *>(X + (size_t)kFeatStride * row);
            const float4 a0 = r[2 * g], a1 = r[2 * g + 1];
            x0[0] = a0.x; x0[1] = a0.y; x0[2] = a0.z; x0[3] = a0.w;
            x0[4] = a1.x; x0[5] = a1.y; x0[6] = a1.z; x0[7] = a1.w;
            if (g < 2) {
                const float4 b0 = r[8 + 2 * g], b1 = r[9 + 2 * g];
                x1[0] = b0.x; x1[1] = b0.y; x1[2] = b0.z; x1[3] = b0.w;
                x1[4] = b1.x; x1[5] = b1.y; x1[6] = b1.z; x1[7] = b1.w;
            }
        }
        x[0][nt] = pack8(x0);
        x[1][nt] = pack8(x1);
    }
}

// Forward over feature rows (stride kFeatStride), raw outputs (n x 3).
__global__ __launch_bounds__(kInferBlock) void forward_rows_kernel(const uint4 *__restrict__ fw,
                                                                   const float *__restrict__ bias,
                                                                   const float *__restrict__ X, uint64_t n,
                                                                   float *__restrict__ out) {
    __shared__ uint4 lds_frag[kFwdFrags * 64];
    __shared__ float lds_bias[kBiasCount];
    stage_weights(fw, bias, lds_frag, lds_bias);
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
    const uint64_t waves = (uint64_t)gridDim.x * (kInferBlock / 64);
    const uint64_t wave = (uint64_t)blockIdx.x * (kInferBlock / 64) + threadIdx.x / 64;
    for (uint64_t base = wave * 32; base < n; base += waves * 32) {
        asm volatile("" ::: "memory");
        uint4 x[2][2];
        load_rows(X, nullptr, base, n, x);
        Acts a;
        f32x4 out2[2];
        forward_wave(lds_frag, lds_bias, x, a, out2);
        if (g == 0) {
            write_out(out2[0], base + p, n, false, out);
            write_out(out2[1], base + 16 + p, n, false, out);
        }
    }
}

// ---- training --------------------------------------------------------------------
// Per-batch buffers, [unit][pair] bf16 with row stride = padded batch size:
// XbT (64 rows), H1T..H3T (128), D4T (16, rows 3.. stay zero), D3T..D1T (128).
struct TrainBufs {
    uint16_t *xb, *h1, *h2, *h3, *d4, *d3, *d2, *d1;
    float *loss;  // per pair
    uint32_t stride;
};

__device__ __forceinline__ void store_tile(uint16_t *T, uint32_t stride, int tm, uint64_t pair0, const float v[4]) {
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) T[(size_t)(16 * tm + 4 * g + i) * stride + pair0 + p] = (uint16_t)f2bf(v[i]);
}

// Stores a B-fragment array (units 32kb + kloc) back to [unit][pair] rows.
__device__ __forceinline__ void store_frags(uint16_t *T, uint32_t stride, const uint4 (*f)[2], int kblocks,
                                            uint64_t pair0) {
#pragma unroll
    for (int kb = 0; kb < kblocks; ++kb) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float lo[4], hi[4];
            for (int i = 0; i < 4; ++i) {
                lo[i] = unpack_slot(f[kb][nt], i);
                hi[i] = unpack_slot(f[kb][nt], 4 + i);
            }
            store_tile(T, stride, 2 * kb, pair0 + 16 * nt, lo);
            store_tile(T, stride, 2 * kb + 1, pair0 + 16 * nt, hi);
        }
    }
}

// Stores the layer-1 B fragments (natural order: slot s of group g in block kb
// is feature 32kb + 8g + s) to [feature][pair] rows.
__device__ __forceinline__ void store_inputs(uint16_t *T, uint32_t stride, const uint4 (*x)[2], uint64_t pair0) {
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const uint4 q = x[kb][nt];
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int s = 0; s < 8; ++s)
                T[(size_t)(32 * kb + 8 * g + s) * stride + pair0 + 16 * nt + p] =
                    (uint16_t)((s & 1) ? (w[s >> 1] >> 16) : (w[s >> 1] & 0xffffu));
        }
    }
}

// dH_in = W^T delta (backward fragments), masked by ReLU'(h_in); result as
// B fragments (for the next backward layer) and stored rows.
template <int KB, class FragPtr>
__device__ __forceinline__ void backward_layer(const FragPtr bw, int frag_base, const uint4 (*delta)[2],
                                               const uint4 (*h_in)[2], uint4 (*d_out)[2]) {
    constexpr int kblocks = KB;
    const uint32_t l = lane_id();
    f32x4 acc[kTiles][2];
#pragma unroll
    for (int tm = 0; tm < kTiles; ++tm) acc[tm][0] = acc[tm][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < kblocks; ++kb) {
#pragma unroll
        for (int tm = 0; tm < kTiles; ++tm) {
            const uint4 a = bw[(frag_base + tm * kblocks + kb) * 64 + l];
            acc[tm][0] = mfma(a, delta[kb][0], acc[tm][0]);
            acc[tm][1] = mfma(a, delta[kb][1], acc[tm][1]);
        }
    }
#pragma unroll
    for (int tm = 0; tm < kTiles; ++tm) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float v[4];
            for (int i = 0; i < 4; ++i) {
                const float h = unpack_slot(h_in[tm >> 1][nt], (tm & 1) * 4 + i);
                v[i] = h > 0.0f ? acc[tm][nt][i] : 0.0f;
            }
            const uint32_t lo = pack2(v[0], v[1]), hi = pack2(v[2], v[3]);
            if (tm & 1) {
                d_out[tm >> 1][nt].z = lo;
                d_out[tm >> 1][nt].w = hi;
            } else {
                d_out[tm >> 1][nt].x = lo;
                d_out[tm >> 1][nt].y = hi;
            }
        }
    }
}

// Forward + loss + back-propagated deltas for one batch (rows idx[0..count)).
__global__ __launch_bounds__(256) void train_fwd_bwd_kernel(const uint4 *__restrict__ fw,
                                                            const uint4 *__restrict__ bw,
                                                            const float *__restrict__ bias,
                                                            const float *__restrict__ X,
                                                            const float *__restrict__ Y,
                                                            const uint32_t *__restrict__ idx, uint32_t count,
                                                            TrainBufs tb) {
    const uint32_t l = lane_id(), g = l >> 4, p = l & 15;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const uint64_t base = wave * 32;
    if (base >= count) return;
    uint4 x[2][2];
    load_rows(X, idx, base, count, x);
    Acts a;
    f32x4 out[2];
    forward_wave(fw, bias, x, a, out);
    store_inputs(tb.xb, tb.stride, x, base);
    store_frags(tb.h1, tb.stride, a.hb1, 4, base);
    store_frags(tb.h2, tb.stride, a.hb2, 4, base);
    store_frags(tb.h3, tb.stride, a.hb3, 4, base);
    // loss and dL/dout (lanes g = 0 own the outputs of pair p)
    uint4 d4[1][2];
    const float inv_b = 1.0f / (float)count;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float dv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint64_t k = base + 16 * nt + p;
        if (g == 0 && k < count) {
            const uint64_t row = idx[k];
            const float y0 = Y[3 * row], y1 = Y[3 * row + 1], y2 = Y[3 * row + 2];
            const float o0 = out[nt][0], o1 = out[nt][1], o2 = out[nt][2];
            const float nn = sqrtf((o0 * o0 + o1 * o1) + o2 * o2);
            const float dot = (o0 * y0 + o1 * y1) + o2 * y2;
            float lossv = 1.0f;
            if (nn > 1e-12f) {
                lossv = 1.0f - dot / nn;
                const float n3 = nn * nn * nn;
                dv[0] = -(y0 / nn - dot * o0 / n3) * inv_b;
                dv[1] = -(y1 / nn - dot * o1 / n3) * inv_b;
                dv[2] = -(y2 / nn - dot * o2 / n3) * inv_b;
            }
            tb.loss[k] = lossv;
        }
        d4[0][nt] = uint4{pack2(dv[0], dv[1]), pack2(dv[2], dv[3]), 0u, 0u};
        if (g == 0) store_tile(tb.d4, tb.stride, 0, base + 16 * nt, dv);
    }
    uint4 d3[4][2], d2[4][2], d1[4][2];
    backward_layer<1>(bw, kBwdL3, d4, a.hb3, d3);
    store_frags(tb.d3, tb.stride, d3, 4, base);
    backward_layer<4>(bw, kBwdL2, d3, a.hb2, d2);
    store_frags(tb.d2, tb.stride, d2, 4, base);
    backward_layer<4>(bw, kBwdL1, d2, a.hb1, d1);
    store_frags(tb.d1, tb.stride, d1, 4, base);
}

// Weight gradients: dW[o][i] = sum over pairs of delta[o][p] * h_in[i][p], one
// 16x16 tile and one chunk of kChunk pairs per wave, K = pairs on the MFMA.
// Bias gradients come from the same delta loads in the ti == 0 tiles.
struct TileTask {
    uint16_t layer, to, ti, pad;
};
__device__ __forceinline__ TileTask task_of(uint32_t t) {
    // layer 0: 8 x 4 tiles; layers 1, 2: 8 x 8; layer 3: 1 x 8
    if (t < 32) return TileTask{0, (uint16_t)(t / 4), (uint16_t)(t % 4), 0};
    t -= 32;
    if (t < 64) return TileTask{1, (uint16_t)(t / 8), (uint16_t)(t % 8), 0};
    t -= 64;
    if (t < 64) return TileTask{2, (uint16_t)(t / 8), (uint16_t)(t % 8), 0};
    t -= 64;
    return TileTask{3, 0, (uint16_t)t, 0};
}

__global__ __launch_bounds__(256) void weight_grad_kernel(TrainBufs tb, uint32_t count, uint32_t chunks,
                                                          float *__restrict__ partial) {
    const uint32_t l = lane_id(), g = l >> 4, r = l & 15;
    const uint32_t task = blockIdx.x * 4 + threadIdx.x / 64;
    if (task >= kTileTasks * chunks) return;
    const uint32_t chunk = task / kTileTasks;
    const TileTask tt = task_of(task % kTileTasks);
    const uint16_t *D = tt.layer == 0 ? tb.d1 : tt.layer == 1 ? tb.d2 : tt.layer == 2 ? tb.d3 : tb.d4;
    const uint16_t *H = tt.layer == 0 ? tb.xb : tt.layer == 1 ? tb.h1 : tt.layer == 2 ? tb.h2 : tb.h3;
    const uint32_t p_begin = chunk * kChunk;
    const uint32_t p_end = min(count, p_begin + kChunk);
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const uint16_t *drow = D + (size_t)(16 * tt.to + r) * tb.stride;
    const uint16_t *hrow = H + (size_t)(16 * tt.ti + r) * tb.stride;
    for (uint32_t p0 = p_begin; p0 < p_end; p0 += 32) {
        // rows are zero beyond `count` up to the padded stride
        const uint4 a = *reinterpret_cast<const uint4 *>(drow + p0 + 8 * g);
        const uint4 b = *reinterpret_cast<const uint4 *>(hrow + p0 + 8 * g);
        acc = mfma(a, b, acc);
        if (tt.ti == 0)
            for (int s = 0; s < 8; ++s) bsum += unpack_slot(a, s);
    }
    float *out = partial + (size_t)chunk * kNumParams;
    const uint32_t in_dim = tt.layer == 0 ? kIn : kH;
    const uint32_t out_dim = tt.layer == 3 ? kOut : kH;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t o = 16 * tt.to + 4 * g + i, in = 16 * tt.ti + r;
        if (o < out_dim && in < in_dim) out[w_offset(tt.layer) + o * in_dim + in] = acc[i];
    }
    if (tt.ti == 0) {
        // lanes r, r+16, r+32, r+48 hold disjoint pair subsets of unit 16to + r
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        const uint32_t o = 16 * tt.to + r;
        if (g == 0 && o < out_dim) out[b_offset(tt.layer) + o] = bsum;
    }
}

// Sums the per-chunk partials in chunk order; either writes the gradient or
// applies one Adam step to the f32 master parameters.
__global__ void reduce_adam_kernel(const float *__restrict__ partial, uint32_t chunks, float *__restrict__ grad_out,
                                   float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2,
                                   AdamStep st) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNumParams) return;
    float gsum = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) gsum += partial[(size_t)c * kNumParams + k];
    if (grad_out) {
        grad_out[k] = gsum;
        return;
    }
    const float a = st.beta1 * m1[k] + (1.0f - st.beta1) * gsum;
    const float b = st.beta2 * m2[k] + (1.0f - st.beta2) * gsum * gsum;
    m1[k] = a;
    m2[k] = b;
    const float ah = a / st.bc1, bh = b / st.bc2;
    params[k] -= st.lr * ah / (sqrtf(bh) + st.eps);
}

// Packs the bf16 fragments (forward and transposed) from the f32 parameters.
__global__ void pack_kernel(const float *__restrict__ params, uint4 *__restrict__ fw, uint4 *__restrict__ bw,
                            float *__restrict__ bias) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = (kFwdFrags + kBwdFrags) * 64;
    if (t < kBiasCount) {
        const uint32_t layer = t < 128 ? 0 : t < 256 ? 1 : t < 384 ? 2 : 3;
        bias[t] = params[b_offset(layer) + (t - 128 * layer)];
    }
    if (t >= total) return;
    const bool fwd = t < kFwdFrags * 64;
    const uint32_t f = (fwd ? t : t - kFwdFrags * 64) / 64, l = t % 64, g = l / 16, r = l % 16;
    float v[8];
    for (int s = 0; s < 8; ++s) {
        uint32_t layer, tm, kb, kblocks;
        if (fwd) {
            if (f < kFragL1) { layer = 0; kblocks = 2; tm = f / 2; kb = f % 2; }
            else if (f < kFragL2) { layer = 1; kblocks = 4; tm = (f - kFragL1) / 4; kb = (f - kFragL1) % 4; }
            else if (f < kFragL3) { layer = 2; kblocks = 4; tm = (f - kFragL2) / 4; kb = (f - kFragL2) % 4; }
            else { layer = 3; kblocks = 4; tm = 0; kb = f - kFragL3; }
            (void)kblocks;
            const uint32_t o = 16 * tm + r;
            const uint32_t k = layer == 0 ? 32 * kb + 8 * g + s : 32 * kb + kloc(g, s);
            const uint32_t in_dim = layer == 0 ? kIn : kH, out_dim = layer == 3 ? kOut : kH;
            v[s] = (o < out_dim && k < in_dim) ? params[w_offset(layer) + o * in_dim + k] : 0.0f;
        } else {
            // transposed: M = the layer's input units, K = its output units
            if (f < kBwdL2) { layer = 3; tm = f; kb = 0; }
            else if (f < kBwdL1) { layer = 2; tm = (f - kBwdL2) / 4; kb = (f - kBwdL2) % 4; }
            else { layer = 1; tm = (f - kBwdL1) / 4; kb = (f - kBwdL1) % 4; }
            const uint32_t in = 16 * tm + r, o = 32 * kb + kloc(g, s);
            const uint32_t out_dim = layer == 3 ? kOut : kH;
            v[s] = o < out_dim ? params[w_offset(layer) + o * kH + in] : 0.0f;
        }
    }
    uint4 q;
    q.x = f2bf(v[0]) | (f2bf(v[1]) << 16);
    q.y = f2bf(v[2]) | (f2bf(v[3]) << 16);
    q.z = f2bf(v[4]) | (f2bf(v[5]) << 16);
    q.w = f2bf(v[6]) | (f2bf(v[7]) << 16);
    if (fwd) fw[f * 64 + l] = q;
    else bw[f * 64 + l] = q;
}

// Sum of the per-pair losses of a batch into a running total (one block, fixed order).
__global__ void loss_sum_kernel(const float *__restrict__ loss, uint32_t count, double *__restrict__ total) {
    __shared__ double part[256];
    double s = 0.0;
    for (uint32_t k = threadIdx.x; k < count; k += 256) s += loss[k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[0] += part[0];
}

// normal1 of world-space contacts, rotated into shape i's frame (GJK labels).
__global__ void local_normals_kernel(const float *__restrict__ contacts, const float *__restrict__ poses,
                                     const uint32_t *__restrict__ pairs, uint64_t m, float *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const float *c = contacts + 14 * k;
    const uint32_t has = __float_as_uint(c[13]);
    f3 n{0.0f, 0.0f, 0.0f};
    if (has == 1u) n = qrot_inv(load_iso(poses + 7 * (size_t)pairs[2 * k]).q, f3{c[6], c[7], c[8]});
    out[3 * k] = n.x;
    out[3 * k + 1] = n.y;
    out[3 * k + 2] = n.z;
}

}  // namespace

// ---- host launchers -----------------------------------------------------------------

hipError_t launch_descriptors(const DeviceShapes &s, float *out, hipStream_t stream) {
    if (s.n == 0) return hipSuccess;
    hipLaunchKernelGGL(descriptors_kernel, dim3((s.n + 255) / 256), dim3(256), 0, stream, s.rec, s.verts, s.n, out);
    return hipGetLastError();
}

hipError_t launch_features(const float *poses, const uint32_t *pairs, uint64_t m, const float *desc, float *out,
                           hipStream_t stream) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3((uint32_t)((m + 255) / 256)), dim3(256), 0, stream, poses, pairs, m, desc,
                       out);
    return hipGetLastError();
}

static hipError_t infer_grid(uint64_t m, uint32_t *blocks) {
    int dev = 0, cus = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const uint64_t per_block = 32ull * (kInferBlock / 64);
    uint64_t want = (m + per_block - 1) / per_block;
    uint64_t b = (uint64_t)cus;  // one block per CU (the weights fill most of the LDS)
    if (b > want) b = want;
    *blocks = (uint32_t)(b < 1 ? 1 : b);
    return hipSuccess;
}

hipError_t launch_seeds(const DeviceMlp &net, const float *poses, const uint32_t *pairs, uint64_t m,
                        const uint64_t *m_dev, const float *desc, float *seeds, hipStream_t stream) {
    if (m == 0) return hipSuccess;
    uint32_t blocks = 0;
    hipError_t e = infer_grid(m, &blocks);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(seeds_kernel, dim3(blocks), dim3(kInferBlock), 0, stream, net.fw, net.bias, poses, pairs, m,
                       (const unsigned long long *)m_dev, desc, seeds);
    return hipGetLastError();
}

hipError_t launch_forward_rows(const DeviceMlp &net, const float *X, uint64_t n, float *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = 0;
    hipError_t e = infer_grid(n, &blocks);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(forward_rows_kernel, dim3(blocks), dim3(kInferBlock), 0, stream, net.fw, net.bias, X, n, out);
    return hipGetLastError();
}

hipError_t launch_pack(const DeviceMlp &net, hipStream_t stream) {
    const uint32_t total = (kFwdFrags + kBwdFrags) * 64;
    hipLaunchKernelGGL(pack_kernel, dim3((total + 255) / 256), dim3(256), 0, stream, net.params, net.fw, net.bw,
                       net.bias);
    return hipGetLastError();
}

hipError_t launch_train_step(const DeviceMlp &net, const TrainWork &w, const float *X, const float *Y,
                             const uint32_t *idx, uint32_t count, const AdamStep *adam, float *grad_out,
                             hipStream_t stream) {
    if (count == 0 || count > w.max_batch) return hipErrorInvalidValue;
    TrainBufs tb{w.xb, w.h1, w.h2, w.h3, w.d4, w.d3, w.d2, w.d1, w.loss, w.max_batch};
    const uint32_t waves = (count + 31) / 32;
    hipLaunchKernelGGL(train_fwd_bwd_kernel, dim3((waves + 3) / 4), dim3(256), 0, stream, net.fw, net.bw, net.bias,
                       X, Y, idx, count, tb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t chunks = (count + kChunk - 1) / kChunk;
    const uint32_t tasks = kTileTasks * chunks;
    hipLaunchKernelGGL(weight_grad_kernel, dim3((tasks + 3) / 4), dim3(256), 0, stream, tb, count, chunks, w.partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const AdamStep st = adam ? *adam : AdamStep{0, 0, 0, 0, 1, 1};
    hipLaunchKernelGGL(reduce_adam_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, stream, w.partial, chunks,
                       grad_out, net.params, net.m1, net.m2, st);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, stream, w.loss, count, w.loss_total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (!grad_out) return launch_pack(net, stream);
    return hipSuccess;
}

hipError_t launch_local_normals(const float *contacts, const float *poses, const uint32_t *pairs, uint64_t m,
                                float *out, hipStream_t stream) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(local_normals_kernel, dim3((uint32_t)((m + 255) / 256)), dim3(256), 0, stream, contacts, poses,
                       pairs, m, out);
    return hipGetLastError();
}

uint16_t host_f2bf(float f) { return (uint16_t)f2bf(f); }

}  // namespace mlp
}  // namespace pq
