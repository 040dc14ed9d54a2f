// shapes.hip -- world-space AABBs of posed shapes (Shape::compute_aabb, src/shape/shape.rs:369).
//
// One lane per shape; the per-type rules are shape_aabb's (contact_gjk.h):
// ConvexPolyhedron: point_cloud_aabb (aabb_utils.rs:66-86) -- every vertex is
// transformed by the pose (quaternion rotation then translation) and folded
// into a running min/max, starting from vertex 0.  Ball: centre -/+ radius
// (aabb_ball.rs:7-27).  Cuboid: centre -/+ |R| * half_extents
// (aabb_cuboid.rs:9-14, utils/isometry_ops.rs:14-19) with R nalgebra's
// quaternion-to-matrix.  Capsule: the posed segment's AABB grown by the radius
// (aabb_capsule.rs:8-20).  Cone: support points along -/+ each axis
// (aabb_support_map.rs:12-14 -> aabb_utils.rs:10-31).  Triangle: the min/max of
// the three posed vertices (aabb_triangle.rs:16-30).
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
    f3 lo, hi;
    shape_aabb(rec[i], verts, load_iso(poses + 7 * (size_t)i), lo, hi);
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
