// shapes.hip -- world-space AABBs of posed shapes (Shape::compute_aabb, src/shape/shape.rs:369).
//
// One lane per shape.  ConvexPolyhedron: point_cloud_aabb (aabb_utils.rs:66-86)
// -- every vertex is transformed by the pose (quaternion rotation then
// translation) and folded into a running min/max, starting from vertex 0.
// Ball: centre -/+ radius (aabb_ball.rs:7-27).  Cuboid: centre -/+
// |R| * half_extents (aabb_cuboid.rs:9-14, utils/isometry_ops.rs:14-19) with R
// nalgebra's quaternion-to-matrix.  Capsule: the posed segment's AABB grown by
// the radius (aabb_capsule.rs:8-20).  Cone: support points along -/+ each axis
// (aabb_support_map.rs:12-14 -> aabb_utils.rs:10-31).
#include "contact_gjk.h"
#include "pq_internal.h"
#include "pq_math.h"

namespace pq {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void shape_aabbs_kernel(const ShapeRecord *__restrict__ rec,
                                                             const float4 *__restrict__ verts, uint32_t n,
                                                             const float *__restrict__ poses,
                                                             float *__restrict__ aabbs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ShapeRecord r = rec[i];
    const iso m = load_iso(poses + 7 * (size_t)i);
    f3 lo, hi;
    if (r.type == PQ_SHAPE_CONVEX_POLYHEDRON) {
        const float4 *v = verts + r.first_vertex;
        float4 p0 = v[0];
        lo = hi = iso_point(m, f3{p0.x, p0.y, p0.z});
        for (uint32_t k = 1; k < r.num_vertices; ++k) {
            const float4 p = v[k];
            const f3 w = iso_point(m, f3{p.x, p.y, p.z});
            lo = min3(lo, w);
            hi = max3(hi, w);
        }
    } else if (r.type == PQ_SHAPE_BALL) {
        const float rad = r.param[0];
        lo = add3(m.t, f3{-rad, -rad, -rad});
        hi = add3(m.t, f3{rad, rad, rad});
    } else if (r.type == PQ_SHAPE_CUBOID) {
        // UnitQuaternion::to_rotation_matrix, then abs(R) * half_extents.
        const float qi = m.q.x, qj = m.q.y, qk = m.q.z, qw = m.q.w;
        const float ww = qw * qw, ii = qi * qi, jj = qj * qj, kk = qk * qk;
        const float ij = qi * qj * 2.0f, wk = qw * qk * 2.0f, wj = qw * qj * 2.0f;
        const float ik = qi * qk * 2.0f, jk = qj * qk * 2.0f, wi = qw * qi * 2.0f;
        const float R[3][3] = {{((ww + ii) - jj) - kk, ij - wk, wj + ik},
                               {wk + ij, ((ww - ii) + jj) - kk, jk - wi},
                               {ik - wj, wi + jk, ((ww - ii) - jj) + kk}};
        const float he[3] = {r.param[0], r.param[1], r.param[2]};
        float w[3];
        for (int a = 0; a < 3; ++a)
            w[a] = (fabsf(R[a][0]) * he[0] + fabsf(R[a][1]) * he[1]) + fabsf(R[a][2]) * he[2];
        lo = sub3(m.t, f3{w[0], w[1], w[2]});
        hi = add3(m.t, f3{w[0], w[1], w[2]});
    } else if (r.type == PQ_SHAPE_CAPSULE) {
        const f3 a = iso_point(m, f3{0.0f, -r.param[0], 0.0f});
        const f3 b = iso_point(m, f3{0.0f, r.param[0], 0.0f});
        const float rad = r.param[1];
        lo = sub3(min3(a, b), f3{rad, rad, rad});
        hi = add3(max3(a, b), f3{rad, rad, rad});
    } else {  // PQ_SHAPE_CONE
        const ShapeRef sr{r.type, 0, nullptr, r.param[0], r.param[1], r.param[2]};
        float mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            f3 basis = f3{0.0f, 0.0f, 0.0f};
            if (d == 0) basis.x = 1.0f;
            if (d == 1) basis.y = 1.0f;
            if (d == 2) basis.z = 1.0f;
            const f3 pmax = iso_point(m, local_support(sr, qrot_inv(m.q, basis)));
            const f3 nb = f3{-basis.x, -basis.y, -basis.z};
            const f3 pmin = iso_point(m, local_support(sr, qrot_inv(m.q, nb)));
            mx[d] = d == 0 ? pmax.x : (d == 1 ? pmax.y : pmax.z);
            mn[d] = d == 0 ? pmin.x : (d == 1 ? pmin.y : pmin.z);
        }
        lo = f3{mn[0], mn[1], mn[2]};
        hi = f3{mx[0], mx[1], mx[2]};
    }
    float *o = aabbs + 6 * (size_t)i;
    o[0] = lo.x;
    o[1] = lo.y;
    o[2] = lo.z;
    o[3] = hi.x;
    o[4] = hi.y;
    o[5] = hi.z;
}

}  // namespace

hipError_t launch_shape_aabbs(const DeviceShapes &s, const float *poses, float *aabbs, hipStream_t stream) {
    if (s.n == 0) return hipSuccess;
    hipLaunchKernelGGL(shape_aabbs_kernel, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, s.rec, s.verts,
                       s.n, poses, aabbs);
    return hipGetLastError();
}

}  // namespace pq
