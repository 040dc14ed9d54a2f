// pq_internal.h -- types shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/parry_mi355x.h"

namespace pq {

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// Validates a pq_ray_tuning (PQ_OK, or PQ_ERR_INVALID with the message set).
int check_ray_tuning(const pq_ray_tuning *t, const char *who);

#define PQ_HIP_CHECK(expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return ::pq::hip_fail(_e, #expr); \
    } while (0)

// ---- Bvh ------------------------------------------------------------------
// Leaf AABB input of the binned builder.
struct LeafBox {
    float mins[3];
    float maxs[3];
};

// Builds parry's binned Bvh (bvh_binned_build.rs:39-176 followed by the
// depth-first relayout of bvh_refit.rs:170-320) over `leaves`.  Output records
// are in depth-first preorder with leaf `children` = leaf index; returns the
// tree depth in wide-node levels through *depth.
void build_binned_bvh(const std::vector<LeafBox> &leaves, std::vector<pq_bvh_node_wide> &out, uint32_t *depth);

// Depth (wide-node levels) of a depth-first tree; 0 if malformed.
uint32_t bvh_depth(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint32_t num_leaves, bool *ok);

inline uint32_t leaf_count(const pq_bvh_node &n) { return n.data & 0x3fffffffu; }
inline bool is_leaf(const pq_bvh_node &n) { return leaf_count(n) == 1; }

// ---- device-side TriMesh layout ------------------------------------------
// Triangle record, 48 B = three 16-B loads: vertices a, b, c and the
// triangle's index in the caller's index buffer.  Records are stored in the
// order in which the depth-first Bvh visits its leaves, so sibling leaves share
// cache lines; Bvh leaf `children` index this array.
struct alignas(16) TriRecord {
    float a[3];
    float b[3];
    float c[3];
    uint32_t tri_id;
    uint32_t pad[2];
};
static_assert(sizeof(TriRecord) == 48, "TriRecord is 48 bytes");
static_assert(sizeof(pq_bvh_node_wide) == 64, "wide node is 64 bytes");

struct DeviceMesh {
    const pq_bvh_node_wide *nodes;  // device, leaf children -> TriRecord index
    const TriRecord *tris;          // device
    uint32_t num_nodes;
    uint32_t num_triangles;
    uint32_t depth;
    uint32_t finite;  // every vertex and every node bound is finite
    pq_ray_tuning tuning;
};

// One mesh of a batch, as the batched kernels see it.
struct BatchMesh {
    const pq_bvh_node_wide *nodes;
    const TriRecord *tris;
    uint32_t ntri;
    uint32_t single_leaf;
    uint32_t num_nodes;
    uint32_t finite;  // DeviceMesh::finite
};

// The binned mapping of a batch sorts each mesh's rays on its own: sort ranges
// of kRayBinRange rays never straddle two meshes, the counts are scanned per
// mesh and the cells' first positions start at the mesh's first ray, so every
// sorted position stays inside its mesh's range of rays.  The buffers belong
// to the batch (allocated when the mapping is first set or tuned, never during
// a graph capture): casts of one batch with this mapping must not overlap.
struct BatchBins {
    uint32_t num_ranges;
    const uint32_t *range_mesh;   // device, per range: its mesh
    const uint32_t *range_first;  // device, num_meshes + 1: each mesh's first range
    uint32_t *hist;               // device, num_ranges x kRayBins: counts, scanned per mesh and cell
    uint32_t *first;              // device, num_meshes x kRayBins: rays per cell, then first sorted positions
    uint32_t *keyrank;            // device, per ray: cell | rank within its range and cell << kRayBinKeyBits
    uint32_t *order;              // device, per sorted position: the ray index (null until allocated)
};

struct DeviceBatch {
    const BatchMesh *meshes;               // device
    const unsigned long long *offsets;     // device, num_meshes + 1 ray offsets
    unsigned long long *queue;             // device, 2 x u64 (persistent mapping)
    uint32_t num_meshes;
    uint32_t max_depth;
    uint64_t num_rays;
    pq_ray_tuning tuning;
    BatchBins bins;
};

hipError_t launch_cast_rays_batch(const DeviceBatch &b, const pq_ray *rays, float max_toi, pq_ray_hit *hits,
                                  uint32_t *counts, hipStream_t stream);

// Per-(device, stream) 16-byte work queue used by persistent kernels
// (zeroed at allocation; every persistent kernel leaves it zeroed).
hipError_t stream_queue(hipStream_t stream, unsigned long long **queue);

// Buffers of the binned ray mapping's counting sort, per (device, stream):
// `hist` (kRayBins x blocks: each sort block's count per cell, scanned per
// cell into offsets), `totals` (kRayBins: rays per cell), `keyrank` (per ray:
// cell | rank within its block and cell << kRayBinKeyBits) and `order` (per
// sorted position: the ray index).  Grown on demand (the stream is
// synchronised before a smaller buffer is freed).
constexpr uint32_t kRayBinBits = 4;  // cells per axis: 16
constexpr uint32_t kRayBinKeyBits = 3 * kRayBinBits;
constexpr uint32_t kRayBins = 1u << kRayBinKeyBits;
constexpr uint32_t kRayBinMaxBlocks = 65536;  // sort blocks; a block ranks < 2^20 rays
constexpr uint32_t kRayBinRange = 4096;       // rays per sort range (before the block cap)
struct RayBinSpace {
    uint32_t *hist;
    uint32_t *totals;
    uint32_t *keyrank;
    uint32_t *order;
};
hipError_t ray_bin_workspace(hipStream_t stream, uint64_t num_rays, uint32_t blocks, RayBinSpace *ws);

// LDS stack entries the ray kernels need for a tree of `depth` wide levels.
uint32_t ray_stack_entries(uint32_t depth);
// Whether the wave-per-ray mapping can run a tree of `depth` wide levels.
bool wave_mapping_fits(uint32_t depth);
// Times every ray mapping candidate with `launch` (which reads `slot`), leaves
// the fastest in `slot` and reports it; synchronises `s`.
// `binned` adds the binned mapping to the candidates.
int autotune_ray_mapping(const std::function<hipError_t()> &launch, pq_ray_tuning &slot, hipStream_t s,
                         pq_ray_tuning *chosen, float *chosen_ms, const char *who, bool binned);

// Launches the ray-cast kernel (ray_trimesh.hip).
// `counts` (nullable, device, 2 per ray): wide-node records fetched, triangles tested.
hipError_t launch_cast_rays(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi, pq_ray_hit *hits,
                            uint32_t *counts, hipStream_t stream);
// The culling variant (1 = ignore back faces, 2 = ignore front faces), static mapping.
// `counts` as for launch_cast_rays.
hipError_t launch_cast_rays_culling(const DeviceMesh &m, const pq_ray *rays, uint64_t n, float max_toi,
                                    uint32_t culling, pq_ray_hit *hits, uint32_t *counts, hipStream_t stream);
// Point projections onto the mesh (point_trimesh.hip); solid as in parry.
// `counts` (nullable, device, 2 per point): wide-node records fetched, triangles tested.
hipError_t launch_project_points(const DeviceMesh &m, const float *points, uint64_t n, int solid,
                                 pq_point_projection *out, uint32_t *counts, hipStream_t stream);

// ---- shapes -------------------------------------------------------------
// Device shape record, 32 B: type tag, vertex range (ConvexPolyhedron) and
// up to four scalar parameters (see pq_shape_desc).
struct alignas(16) ShapeRecord {
    uint32_t type;
    uint32_t first_vertex;
    uint32_t num_vertices;
    uint32_t pad;
    float param[4];
};
static_assert(sizeof(ShapeRecord) == 32, "ShapeRecord is 32 bytes");

struct DeviceShapes {
    const ShapeRecord *rec;  // device
    const float4 *verts;     // device, xyz + pad
    uint32_t n;
    uint32_t num_vertices;
    // ConvexPolyhedron faces (null until pq_shapes_set_polyhedron_faces); a
    // polyhedron's record holds its first face and face count in param[0] and
    // param[1] as uint32 bit patterns.
    const pq_polyhedron_face *faces = nullptr;
    const uint32_t *face_verts = nullptr;
    const uint32_t *face_edges = nullptr;
};

hipError_t launch_shape_aabbs(const DeviceShapes &s, const float *poses, float *aabbs, hipStream_t stream);

// ---- broadphase Bvh -----------------------------------------------------
struct DeviceBvh {
    pq_bvh_node_wide *nodes;  // device, depth-first
    uint32_t *leaf_slot;      // device, per leaf: wide * 2 + (1 if right)
    uint32_t *parent_slot;    // device, per wide node: parent wide * 2 + side (root: UINT32_MAX)
    uint32_t *pair_counts;    // device, n + 1
    unsigned long long *pair_offsets;  // device, n + 1
    void *scan_temp;
    size_t scan_temp_bytes;
    uint32_t num_leaves;
    uint32_t num_nodes;
    uint32_t depth;
    // Refit order: the non-root wide nodes grouped by height above the leaves
    // (a node's children all have smaller heights).
    uint32_t *level_nodes;             // device, the nodes level by level
    uint32_t *level_offsets;           // device, num_levels + 1
    const uint32_t *level_offsets_host;  // the same offsets on the host
    uint32_t num_levels;
};

hipError_t launch_bvh_refit(const DeviceBvh &b, const float *aabbs, hipStream_t stream);
// Bvh::from_leaves(Ploc, aabbs) on the device (ploc.hip): n - 1 depth-first wide
// records (one for n <= 2) into `out`; *depth set on return.  Synchronises `stream`.
size_t ploc_workspace_bytes(uint32_t n);
hipError_t build_ploc_bvh(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t *depth, void *workspace,
                          hipStream_t stream);
hipError_t launch_bvh_self_pairs(const DeviceBvh &b, uint32_t *pairs, uint64_t capacity, unsigned long long *count,
                                 hipStream_t stream);
hipError_t bvh_scan_temp_bytes(uint32_t num_leaves, size_t *bytes);

// ---- contacts ------------------------------------------------------------
// Waves (and 438 KB global-memory polytope buffers) of the EPA pass that
// re-runs pairs whose polytope outgrew the LDS capacity.
constexpr uint32_t kEpaGlobalWaves = 32;
size_t contact_workspace_bytes(uint64_t num_pairs);

// TriMesh-against-shapes contacts (composite.hip).
hipError_t build_mesh_shape_records(const DeviceMesh &m, ShapeRecord *rec, float4 *verts, hipStream_t stream);
size_t mesh_contact_workspace_bytes(uint32_t ntri, uint32_t n, uint64_t capacity);
hipError_t launch_mesh_contacts(const DeviceMesh &m, const DeviceShapes &combined, uint32_t n, const float *mesh_pose,
                                const float *poses, float prediction, uint64_t capacity, pq_contact *out,
                                uint32_t *triangles, unsigned long long *num_candidates, void *workspace,
                                hipStream_t stream);
// m_dev (nullable, device): the number of pairs actually present (<= m), e.g.
// the count written by pq_bvh_self_pairs, so a frame needs no host round trip.
// flags: kContactLocalSupportMaps runs contact_support_map_support_map on every
// pair (Ball included, no closed forms) and leaves the results in the shapes'
// local frames, as the shape cast's impact geometry needs.
constexpr uint32_t kContactLocalSupportMaps = 1u;
// kContactFeatureMaps runs it on the pairs' polygonal feature maps (a Capsule is its
// segment) with prediction + both border radii, local frames, and keeps GJK's
// separating direction in normal1 of a pair without contact; a pair without
// feature maps gets PQ_CONTACT_UNSUPPORTED.
constexpr uint32_t kContactFeatureMaps = 2u;
// kContactShortEpa: the pairs that reach EPA need few iterations each (a Triangle
// against a shape), so the first EPA tier's automatic schedule takes them per
// wave rather than refilling each group (the set-up of a pair would then run
// beside the other groups' iterations about once per iteration).
constexpr uint32_t kContactShortEpa = 4u;
hipError_t launch_contacts(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                           const uint64_t *m_dev, float prediction, const float *seeds, pq_contact *out,
                           uint32_t *iters, void *workspace, hipStream_t stream, uint32_t flags = 0u);
// Contact manifolds between polygonal feature maps (manifold.hip).
size_t manifold_workspace_bytes(uint64_t m);
hipError_t launch_manifolds_update(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                                   float prediction, pq_contact_manifold *manifolds, void *workspace,
                                   hipStream_t stream);
hipError_t launch_manifolds(const DeviceShapes &s, const float *poses, const uint32_t *pairs, uint64_t m,
                            float prediction, pq_contact_manifold *out, void *workspace, hipStream_t stream);
// Linear shape casts (cast.hip).
size_t cast_workspace_bytes(uint64_t m);
hipError_t launch_casts(const DeviceShapes &s, const float *poses, const float *vels, const uint32_t *pairs, uint64_t m,
                        const pq_shape_cast_options &o, pq_shape_cast_hit *out, uint32_t *iters, void *workspace,
                        hipStream_t stream);
// Lanes that run one pair's EPA together (64, 16 or 8), set by pq_contact_set_tuning.
extern int g_epa_lanes_per_pair;
extern int g_epa_polytope;
extern int g_epa_schedule;  // first EPA tier: 0 = automatic, 1 = per wave, 2 = a refill per group
// Phase timing of launch_contacts (pq_contact_set_phase_timing): while on, each
// call records kContactPhases + 1 events on its stream around its phases.
constexpr int kContactPhases = 5;  // GJK/dispatch, first EPA tier, 768-face, 2048-face, global-memory tier
extern bool g_contact_phase_timing;
extern hipEvent_t g_contact_phase_events[kContactPhases + 1];
// Per-(device, stream) workspace owned by the library, grown on demand.
hipError_t stream_workspace(hipStream_t stream, size_t bytes, void **ptr);

}  // namespace pq

struct pq_shapes {
    pq::DeviceShapes dev;
    std::vector<pq::ShapeRecord> host_rec;
    void *d_rec = nullptr;
    void *d_verts = nullptr;
    void *d_faces = nullptr;  // faces, face vertices, face edges (one allocation)
};

struct pq_bvh {
    pq::DeviceBvh dev;
    std::vector<void *> allocations;
    std::vector<uint32_t> level_offsets;  // host copy backing dev.level_offsets_host
};

struct pq_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct pq_trimesh_batch {
    pq::DeviceBatch dev;
    std::vector<void *> allocations;
    std::vector<uint64_t> offsets;  // host copy of the ray offsets
};

struct pq_mesh_contact {
    const pq_trimesh *mesh = nullptr;
    pq::DeviceShapes combined{};  // the mesh's triangles (Bvh leaf record order), then the shapes
    uint32_t num_shapes = 0;
    void *d_rec = nullptr;
    void *d_verts = nullptr;
};

struct pq_trimesh {
    pq::DeviceMesh dev;
    std::vector<pq_bvh_node_wide> host_nodes;  // leaf children = triangle index
    void *d_nodes = nullptr;
    void *d_tris = nullptr;
};
