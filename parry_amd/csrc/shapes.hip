// shapes.hip -- world-space AABBs of posed shapes (Shape::compute_aabb, src/shape/shape.rs:369).
//
// One lane per shape.  ConvexPolyhedron: point_cloud_aabb (aabb_utils.rs:66-86)
// -- every vertex is transformed by the pose (quaternion rotation then
// translation) and folded into a running min/max, starting from vertex 0.
// Ball: centre -/+ radius (aabb_ball.rs:7-27).
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
    } else {  // PQ_SHAPE_BALL (other types are rejected at upload)
        const float rad = r.param[0];
        lo = add3(m.t, f3{-rad, -rad, -rad});
        hi = add3(m.t, f3{rad, rad, rad});
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
