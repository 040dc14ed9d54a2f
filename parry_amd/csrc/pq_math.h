// pq_math.h -- f32 vector and isometry arithmetic for the gfx950 kernels.
//
// Every helper evaluates parry3d's (nalgebra 0.34's) f32 expression in the same
// order, because the kernels' outputs are compared bit for bit with parry:
//   dot(a, b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a, b) = (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx)
//   q * v       = (t*q.w + q.v x t) + v  with  t = (q.v x v) * 2
//   iso * p     = q * p + translation
// Translation units compiled with -ffp-contract=off keep these sequences free
// of fused multiply-adds.
#pragma once

#include <hip/hip_runtime.h>

namespace pq {

struct f3 {
    float x, y, z;
};

__host__ __device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 neg3(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ f3 mul3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ f3 div3(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ float norm_sq3(f3 a) { return dot3(a, a); }
__host__ __device__ __forceinline__ f3 min3(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__host__ __device__ __forceinline__ f3 max3(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

struct quat {
    float x, y, z, w;  // [i, j, k, w]
};

struct iso {
    quat q;
    f3 t;
};

__host__ __device__ __forceinline__ f3 qrot(const quat &q, f3 v) {
    const f3 qv = f3{q.x, q.y, q.z};
    const f3 t = mul3(cross3(qv, v), 2.0f);
    const f3 c = cross3(qv, t);
    return add3(add3(mul3(t, q.w), c), v);
}
__host__ __device__ __forceinline__ quat qconj(const quat &q) { return quat{-q.x, -q.y, -q.z, q.w}; }
__host__ __device__ __forceinline__ f3 qrot_inv(const quat &q, f3 v) { return qrot(qconj(q), v); }
__host__ __device__ __forceinline__ quat qmul(const quat &a, const quat &b) {
    quat r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return r;
}
__host__ __device__ __forceinline__ f3 iso_point(const iso &m, f3 p) { return add3(qrot(m.q, p), m.t); }
// Isometry::inv_mul: (q_a^-1 * q_b, q_a^-1 * (t_b - t_a)).
__host__ __device__ __forceinline__ iso iso_inv_mul(const iso &a, const iso &b) {
    const quat ia = qconj(a.q);
    return iso{qmul(ia, b.q), qrot(ia, sub3(b.t, a.t))};
}

// Isometry::inverse: (q^-1, q^-1 * -t).
__host__ __device__ __forceinline__ iso iso_inverse(const iso &m) {
    const quat qi = qconj(m.q);
    return iso{qi, qrot(qi, neg3(m.t))};
}
__device__ __forceinline__ iso load_iso(const float *p) {
    return iso{quat{p[0], p[1], p[2], p[3]}, f3{p[4], p[5], p[6]}};
}

}  // namespace pq
