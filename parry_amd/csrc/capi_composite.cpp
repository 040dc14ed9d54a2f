// capi_composite.cpp -- C ABI of the contacts between a TriMesh and convex shapes.
#include <cstdint>
#include <string>
#include <vector>

#include "pq_internal.h"

using namespace pq;

extern "C" {

int pq_mesh_contact_destroy(pq_mesh_contact *mc) {
    if (!mc) return PQ_OK;
    if (mc->d_rec) (void)hipFree(mc->d_rec);
    if (mc->d_verts) (void)hipFree(mc->d_verts);
    delete mc;
    return PQ_OK;
}

int pq_mesh_contact_create(const pq_trimesh *mesh, const pq_shapes *shapes, pq_mesh_contact **out) {
    if (!out) return fail(PQ_ERR_INVALID, "pq_mesh_contact_create: null output");
    *out = nullptr;
    if (!mesh || !shapes) return fail(PQ_ERR_INVALID, "pq_mesh_contact_create: null mesh or shapes");
    const uint32_t ntri = mesh->dev.num_triangles, n = shapes->dev.n;
    const uint64_t nrec = (uint64_t)ntri + n, nverts = 3ull * ntri + shapes->dev.num_vertices;
    if (nrec > UINT32_MAX || nverts > UINT32_MAX)
        return fail(PQ_ERR_INVALID, "pq_mesh_contact_create: more than 2^32 records or vertices");
    auto *mc = new pq_mesh_contact();
    hipError_t e = hipMalloc(&mc->d_rec, nrec * sizeof(ShapeRecord) + 16);
    if (e == hipSuccess) e = hipMalloc(&mc->d_verts, nverts * sizeof(float4) + 16);
    if (e == hipSuccess) e = build_mesh_shape_records(mesh->dev, (ShapeRecord *)mc->d_rec, (float4 *)mc->d_verts, nullptr);
    // The shapes follow the triangles; their vertex ranges move past the triangles' vertices.
    std::vector<ShapeRecord> rec = shapes->host_rec;
    for (ShapeRecord &r : rec) r.first_vertex += 3 * ntri;
    if (e == hipSuccess && n)
        e = hipMemcpy((ShapeRecord *)mc->d_rec + ntri, rec.data(), n * sizeof(ShapeRecord), hipMemcpyHostToDevice);
    if (e == hipSuccess && shapes->dev.num_vertices)
        e = hipMemcpy((float4 *)mc->d_verts + 3ull * ntri, shapes->d_verts, shapes->dev.num_vertices * sizeof(float4),
                      hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        pq_mesh_contact_destroy(mc);
        return hip_fail(e, "pq_mesh_contact_create");
    }
    mc->mesh = mesh;
    mc->num_shapes = n;
    mc->combined = DeviceShapes{(const ShapeRecord *)mc->d_rec, (const float4 *)mc->d_verts, (uint32_t)nrec,
                                (uint32_t)nverts};
    *out = mc;
    return PQ_OK;
}

size_t pq_mesh_contact_workspace_bytes(const pq_mesh_contact *mc, uint64_t max_candidates) {
    if (!mc) return 0;
    return mesh_contact_workspace_bytes(mc->mesh->dev.num_triangles, mc->num_shapes, max_candidates);
}

int pq_mesh_contact_shapes(const pq_mesh_contact *mc, const float *mesh_pose, const float *poses, float prediction,
                           uint64_t max_candidates, pq_contact *out, uint32_t *triangles, uint64_t *num_candidates,
                           void *workspace, void *stream) {
    if (!mc) return fail(PQ_ERR_INVALID, "pq_mesh_contact_shapes: null handle");
    if (mc->num_shapes == 0) return PQ_OK;
    if (!mesh_pose || !poses || !out || !workspace)
        return fail(PQ_ERR_INVALID, "pq_mesh_contact_shapes: null mesh pose, poses, output or workspace");
    PQ_HIP_CHECK(launch_mesh_contacts(mc->mesh->dev, mc->combined, mc->num_shapes, mesh_pose, poses, prediction,
                                      max_candidates, out, triangles, (unsigned long long *)num_candidates,
                                      workspace, (hipStream_t)stream));
    return PQ_OK;
}

}  // extern "C"
