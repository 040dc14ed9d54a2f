/*
 * parry_mi355x.h -- C ABI of the MI355X (gfx950) batched collision-query path.
 *
 * Plain C: opaque handles, plain pointers and sizes, no HIP or torch types.
 * Every entry point returns an int status (PQ_OK on success) and, on failure,
 * records a message retrievable with pq_last_error() (thread-local).
 * Per-query "no result" is a sentinel, mirroring parry's Option:
 *   ray hit:  feature == PQ_NO_HIT  <=> None
 * Device-pointer entry points take buffers in GPU memory (allocated with
 * pq_device_malloc or by the caller's own runtime) and an optional HIP stream
 * passed as `void*` (NULL = the default stream); they are asynchronous with
 * respect to the host.  `*_host` variants copy in and out and synchronise.
 *
 * Citations are to dimforge/parry 0.25.3 (path:line).
 */
#ifndef PARRY_MI355X_H
#define PARRY_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PQ_OK 0
#define PQ_ERR_INVALID 1     /* bad argument (null pointer, size, index) */
#define PQ_ERR_UNSUPPORTED 2 /* parry's Unsupported, src/query/error.rs:146-147 */
#define PQ_ERR_OOM 3         /* device allocation failed */
#define PQ_ERR_HIP 4         /* HIP runtime error (message has the detail) */
#define PQ_ERR_CAPACITY 5    /* an output buffer was too small; required size reported */

#define PQ_NO_HIT 0xFFFFFFFFu

/* Copies the calling thread's last error message (NUL-terminated, truncated
 * to `cap`).  Returns the full message length. */
size_t pq_last_error(char *buf, size_t cap);

/* ---- device memory helpers (for callers without HIP bindings) ---------- */
int pq_device_malloc(void **ptr, size_t bytes);
int pq_device_free(void *ptr);
int pq_memcpy_htod(void *dst_device, const void *src_host, size_t bytes, void *stream);
int pq_memcpy_dtoh(void *dst_host, const void *src_device, size_t bytes, void *stream);
int pq_stream_synchronize(void *stream);

/* ---- geometry records --------------------------------------------------- */

/* parry3d::query::Ray (src/query/ray/ray.rs:75-88, #[repr(C)], f32): origin then
 * direction; the direction need not be unit length.  A Rust `&[Ray]` can be
 * passed as `const pq_ray*` without conversion. */
typedef struct pq_ray {
    float origin[3];
    float dir[3];
} pq_ray;

/* parry3d::query::RayIntersection (src/query/ray/ray.rs:286-315) for a TriMesh:
 * `feature` is the index k of FeatureId::Face(k) as remapped by
 * src/query/ray/ray_trimesh.rs:25-33 (k < num_triangles: front face of
 * triangle k; k >= num_triangles: back face of triangle k - num_triangles),
 * or PQ_NO_HIT for None. */
typedef struct pq_ray_hit {
    float time_of_impact;
    float normal[3];
    uint32_t feature;
} pq_ray_hit;

/* parry3d BvhNodeWide (src/partitioning/bvh/bvh_tree.rs:263-266) in f32: two
 * BvhNode records (bvh_tree.rs:456-467) of {mins, children, maxs, data}.  For
 * a leaf, `children` is the leaf (triangle) index; otherwise the index of the
 * child wide node.  `data` holds leaf_count in its low 30 bits (a node is a
 * leaf iff leaf_count == 1; a right child with leaf_count == 0 marks a root
 * with a single leaf) and the change flags in the top 2 bits. */
typedef struct pq_bvh_node {
    float mins[3];
    uint32_t children;
    float maxs[3];
    uint32_t data;
} pq_bvh_node;
typedef struct pq_bvh_node_wide {
    pq_bvh_node left;
    pq_bvh_node right;
} pq_bvh_node_wide;

/* ---- Bvh construction (host only, no GPU needed) ---------------------------- */

/* Bvh::from_leaves(BvhBuildStrategy::Binned, aabbs) (src/partitioning/bvh/bvh_tree.rs:1835-1837,
 * 1891-1955): parry's binned top-down build (bvh_binned_build.rs:39-176) with the
 * depth-first layout produced by its refit (bvh_refit.rs:170-320).
 *   aabbs: n records of {mins[3], maxs[3]} (f32); leaf i gets `children` = i.
 * Writes *num_nodes wide records (max(1, n-1) for n >= 1; 0 for n == 0) and the
 * depth in wide-node levels.  PQ_ERR_CAPACITY (with *num_nodes set) if
 * `capacity` is too small. */
int pq_bvh_from_leaves(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t capacity,
                       uint32_t *num_nodes, uint32_t *depth);

/* Bvh::from_leaves(BvhBuildStrategy::Ploc, aabbs) (bvh_tree.rs:1891-1955,
 * bvh_ploc_build.rs:10-93, utils/morton.rs:13-40) built on the GPU: Morton
 * keys, a stable radix sort, the PLOC merge levels and refit's depth-first
 * layout, identical bit for bit to parry's tree.
 *   aabbs: device, n x {mins[3], maxs[3]}; out: device, `capacity` wide records
 *   (n - 1 needed; 1 for n <= 2); workspace: device,
 *   pq_bvh_ploc_workspace_bytes(n) bytes, or NULL for a library-held one.
 * Synchronises `stream` (the host sizes each merge level); *num_nodes and
 * *depth (nullable) are set on return. */
size_t pq_bvh_ploc_workspace_bytes(uint32_t n);
int pq_bvh_from_leaves_ploc(const float *aabbs, uint32_t n, pq_bvh_node_wide *out, uint32_t capacity,
                            uint32_t *num_nodes, uint32_t *depth, void *workspace, void *stream);

/* ---- Serialized Bvh and TriMesh (host only, no GPU needed) --------------------
 * parry derives serde for Bvh {nodes, parents, leaf_node_indices}
 * (bvh_tree.rs:1748-1762; BvhNodeWide, BvhNode, BvhNodeData, BvhNodeIndex at
 * bvh_tree.rs:148-156, 252-266, 447-467, 1243-1250; VecMap {n, v},
 * utils/vec_map.rs:73-88) and TriMesh {bvh, vertices, indices, pseudo_normals,
 * topology, connected_components, flags} (trimesh.rs:510-528).  These calls
 * read and write those fields as bincode 1.x's default options lay them out
 * (little-endian, fixed-width integers, usize as u64, Vec = u64 length then the
 * elements, Option = u8 tag then the value, structs and fixed arrays with
 * nothing in between), so a node is its 64-byte record and a BvhNodeIndex the
 * u64 (wide << 1) | is_right.  The encoders derive parents and
 * leaf_node_indices from depth-first nodes as Bvh::from_leaves and refit leave
 * them (parents[0] = 0; v[leaf] = Some(slot) for leaves 0..n-1); the decoders
 * check them against the nodes.  TriMesh: the mesh TriMesh::new makes
 * (pseudo-normals, topology and connected components None, flags empty;
 * trimesh.rs:607-612, 718-746); the decoder returns PQ_ERR_UNSUPPORTED for a
 * mesh with any of them set.  *size / counts are set also on PQ_ERR_CAPACITY
 * (out == NULL queries the size). */
int pq_bvh_encode(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint8_t *out, size_t capacity, size_t *size);
int pq_bvh_decode(const uint8_t *bytes, size_t size, pq_bvh_node_wide *nodes, uint32_t capacity, uint32_t *num_nodes);
int pq_trimesh_encode(const float *vertices, uint32_t num_vertices, const uint32_t *indices, uint32_t num_triangles,
                      const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint8_t *out, size_t capacity,
                      size_t *size);
int pq_trimesh_decode(const uint8_t *bytes, size_t size, float *vertices, uint32_t vertex_capacity,
                      uint32_t *num_vertices, uint32_t *indices, uint32_t triangle_capacity, uint32_t *num_triangles,
                      pq_bvh_node_wide *nodes, uint32_t node_capacity, uint32_t *num_nodes);

/* ---- TriMesh ray casts --------------------------------------------------- */

typedef struct pq_trimesh pq_trimesh;

/* Creates a device-resident TriMesh (parry: TriMesh::new, src/shape/trimesh.rs:607-612,
 * which builds its Bvh with the binned strategy, trimesh.rs:1159-1171).
 *   vertices: 3*num_vertices floats; indices: 3*num_triangles u32.
 *   nodes:    NULL to build the Bvh here with parry's binned builder
 *             (bvh_binned_build.rs:39-176 + depth-first refit, bvh_refit.rs:170-320);
 *             or the caller's Bvh exported as depth-first wide nodes
 *             (e.g. from a Rust Bvh), num_nodes records.
 * Host pointers; the data is copied.  Fails with PQ_ERR_INVALID for an empty
 * index buffer (TriMeshBuilderError::EmptyIndices, trimesh.rs:723-725) or an
 * out-of-range vertex index. */
int pq_trimesh_create(const float *vertices, uint32_t num_vertices, const uint32_t *indices,
                      uint32_t num_triangles, const pq_bvh_node_wide *nodes, uint32_t num_nodes,
                      pq_trimesh **out);
int pq_trimesh_destroy(pq_trimesh *mesh);
/* Number of Bvh wide nodes, and the tree depth in wide-node levels. */
int pq_trimesh_info(const pq_trimesh *mesh, uint32_t *num_nodes, uint32_t *depth, uint32_t *num_triangles);
/* Copies the mesh's Bvh (depth-first wide nodes, leaf `children` = triangle index) to host memory. */
int pq_trimesh_export_bvh(const pq_trimesh *mesh, pq_bvh_node_wide *out_host, uint32_t capacity);

/* Kernel mapping of the ray casts, chosen per mesh or per batch (no
 * counterpart in parry; results, and the per-ray node and triangle counts, are
 * identical for every setting, only speed changes):
 *   PQ_RAY_MAPPING_STATIC      one launch lane per ray;
 *   PQ_RAY_MAPPING_PERSISTENT  resident waves pull `chunk` ray indices at a time
 *                              from a work queue in device memory and refill
 *                              finished lanes every iteration;
 *   PQ_RAY_MAPPING_WAVE        one wave per ray: the wave fetches 64 consecutive
 *                              wide nodes at once (lane k: node base + k, with
 *                              the slab tests of both children and the triangle
 *                              tests of its leaves against max_toi), then walks
 *                              find_best's order over them, refetching when the
 *                              walk leaves the window; waves pull `chunk` rays at
 *                              a time from the queue.  For few rays on a large
 *                              mesh.  Trees deeper than 65 wide levels run the
 *                              persistent mapping instead;
 *   PQ_RAY_MAPPING_BINNED      the static mapping over the rays in the order of
 *                              the cell (16^3 grid over the Bvh's root box) where
 *                              each enters that box: a counting sort by cell
 *                              runs first, so the rays of a wave enter the mesh
 *                              near each other and take similar paths; hits
 *                              still land at the rays' own indices.  For large
 *                              batches of incoherent rays; at most 2^32 - 1 rays
 *                              per call.  In a batch each mesh's rays are sorted
 *                              on their own (cells of that mesh's root box) into
 *                              sort buffers the batch owns: 8 bytes per ray and
 *                              16 KiB per 4096 rays of a mesh, allocated when the
 *                              mapping is first set or tuned (at most 65535
 *                              meshes and 65536 sort ranges).
 * block_size (static, persistent, binned): lanes per workgroup, 64, 128 or 256.
 * stack_regs (static, persistent, binned): find_best's pending-node stack keeps
 * its top 0, 1, 2 or 4 entries in registers and the rest in LDS.
 * The queue (persistent and wave) and the sort's buffers (binned, grown to the
 * largest call) of a stream are allocated on the first such call on that
 * stream: make that first call outside hipGraph capture. */
#define PQ_RAY_MAPPING_STATIC 0u
#define PQ_RAY_MAPPING_PERSISTENT 1u
#define PQ_RAY_MAPPING_WAVE 2u
#define PQ_RAY_MAPPING_BINNED 3u
typedef struct pq_ray_tuning {
    uint32_t mapping;       /* PQ_RAY_MAPPING_* */
    uint32_t chunk;         /* persistent and wave: rays claimed per queue access (0 = 128; wave: 0 = 4) */
    uint32_t blocks_per_cu; /* persistent and wave: resident workgroups per CU (0 = occupancy limit) */
    uint32_t block_size;    /* static, persistent, binned: 64, 128 or 256 lanes (0 = 256) */
    uint32_t stack_regs;    /* static, persistent, binned: 0, 1, 2 or 4 stack entries in registers */
    uint32_t reserved;      /* must be 0 */
} pq_ray_tuning;
int pq_trimesh_set_tuning(pq_trimesh *mesh, const pq_ray_tuning *tuning);
int pq_trimesh_get_tuning(const pq_trimesh *mesh, pq_ray_tuning *tuning);
/* Times the kernel mappings (as pq_trimesh_batch_autotune, and the binned
 * mapping over block sizes and 0 or 2 register stack entries) on these rays
 * (device, num_rays > 0; hits receives their results), keeps the fastest in the
 * mesh's tuning and returns it with its time per launch.  Synchronises
 * `stream`.  The best mapping depends on the number of rays.  The culled casts
 * (pq_trimesh_cast_local_rays_with_culling) follow the binned mapping when it
 * is selected and run static lanes otherwise. */
int pq_trimesh_autotune(pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays, float max_toi, pq_ray_hit *hits,
                        void *stream, pq_ray_tuning *chosen, float *chosen_ms);

/* Batched RayCast::cast_local_ray_and_get_normal on a TriMesh
 * (src/query/ray/ray.rs:393-403 -> ray_trimesh.rs:16-35 -> Bvh::find_best,
 * bvh_traverse.rs:335-417).  One result per ray; bit-identical triangle/feature
 * choice to parry3d (f32).  Device pointers; asynchronous on `stream`. */
int pq_trimesh_cast_local_rays(const pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays, float max_toi,
                               int solid, pq_ray_hit *hits, void *stream);
/* Instrumented variant (no counterpart in parry): also writes, per ray, two
 * u32 counts to `counts` (device, 2*num_rays): the number of Bvh wide-node
 * records fetched (iterations of find_best's loop, bvh_traverse.rs:361-416)
 * and of triangles tested.  Used for traffic accounting and to prove the
 * traversal identical to parry's, not for throughput. */
int pq_trimesh_cast_local_rays_counted(const pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays, float max_toi,
                                       int solid, pq_ray_hit *hits, uint32_t *counts, void *stream);
/* Same with host buffers (copies in/out, synchronous). */
int pq_trimesh_cast_local_rays_host(const pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays, float max_toi,
                                    int solid, pq_ray_hit *hits);

/* Batched TriMesh::cast_local_ray_with_culling (src/query/ray/ray_trimesh.rs:152-178):
 * the same traversal, skipping the triangles RayCullingMode::check rejects
 * (ray_trimesh.rs:58-65), with the same feature remap.  Device pointers;
 * asynchronous on `stream`. */
#define PQ_CULL_BACKFACES 1u  /* RayCullingMode::IgnoreBackfaces */
#define PQ_CULL_FRONTFACES 2u /* RayCullingMode::IgnoreFrontfaces */
int pq_trimesh_cast_local_rays_with_culling(const pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays,
                                            float max_toi, uint32_t culling, pq_ray_hit *hits, void *stream);
/* Instrumented variant (no counterpart in parry), counts as for
 * pq_trimesh_cast_local_rays_counted; a culled triangle counts as tested. */
int pq_trimesh_cast_local_rays_with_culling_counted(const pq_trimesh *mesh, const pq_ray *rays, uint64_t num_rays,
                                                    float max_toi, uint32_t culling, pq_ray_hit *hits,
                                                    uint32_t *counts, void *stream);

/* PointProjection (src/query/point/point_query.rs:65-77) plus the index of
 * the triangle projected on: FeatureId::Face(triangle) of
 * TriMesh::project_local_point_and_get_feature (point_composite_shape.rs:171-186,
 * a mesh without pseudo-normals). */
typedef struct pq_point_projection {
    float point[3];
    uint32_t is_inside;
    uint32_t triangle;
} pq_point_projection;

/* Batched PointQuery::project_local_point on a TriMesh
 * (point_composite_shape.rs:164-168 -> CompositeShapeRef::project_local_point,
 * :59-80 -> Bvh::find_best with the AABB distance as node cost, bvh_queries.rs:213-231).
 * points: device, 3 floats per point.  solid as in parry (it changes the
 * result only for points whose projection falls on a degenerate triangle's
 * face region).  With solid = 0 the triangle is FeatureId::Face of
 * project_local_point_and_get_feature.  Device pointers; asynchronous. */
int pq_trimesh_project_local_points(const pq_trimesh *mesh, const float *points, uint64_t num_points, int solid,
                                    pq_point_projection *out, void *stream);
/* Instrumented variant (no counterpart in parry): also writes, per point, two
 * u32 counts to `counts` (device, 2*num_points): the Bvh wide-node records
 * fetched (iterations of find_best's loop) and the triangles tested. */
int pq_trimesh_project_local_points_counted(const pq_trimesh *mesh, const float *points, uint64_t num_points,
                                            int solid, pq_point_projection *out, uint32_t *counts, void *stream);

/* ---- Batched ray casts over many TriMeshes (no counterpart in parry: the
 *      batching of many RayCast calls) --------------------------------------- */

typedef struct pq_trimesh_batch pq_trimesh_batch;
typedef struct pq_graph pq_graph;

/* Groups meshes for one launch.  ray_offsets (host, num_meshes + 1, first 0,
 * non-decreasing): the rays of mesh i are [ray_offsets[i], ray_offsets[i+1]) of
 * the concatenated ray array passed to the cast calls.  The meshes must
 * outlive the batch. */
int pq_trimesh_batch_create(const pq_trimesh *const *meshes, uint32_t num_meshes, const uint64_t *ray_offsets,
                            pq_trimesh_batch **out);
int pq_trimesh_batch_destroy(pq_trimesh_batch *batch);
int pq_trimesh_batch_set_tuning(pq_trimesh_batch *batch, const pq_ray_tuning *tuning);
int pq_trimesh_batch_get_tuning(const pq_trimesh_batch *batch, pq_ray_tuning *tuning);
/* Every ray of every mesh: per-ray results exactly as pq_trimesh_cast_local_rays
 * on the ray's own mesh.  Device pointers; asynchronous on `stream`.  The
 * persistent mapping uses a work queue owned by the batch, the binned mapping
 * the batch's sort buffers: launches on one batch must be ordered (one stream
 * or a graph). */
int pq_trimesh_batch_cast_local_rays(const pq_trimesh_batch *batch, const pq_ray *rays, float max_toi, int solid,
                                     pq_ray_hit *hits, void *stream);
int pq_trimesh_batch_cast_local_rays_counted(const pq_trimesh_batch *batch, const pq_ray *rays, float max_toi,
                                             int solid, pq_ray_hit *hits, uint32_t *counts, void *stream);
/* Times the kernel mappings on these rays (static and persistent lanes over
 * block sizes and register stack depths, persistent chunk sizes and
 * residencies, wave per ray, and the binned mapping when the batch fits its
 * sort, whose buffers this then allocates), keeps the fastest in the batch and
 * returns it with its time per launch.  Synchronises `stream`. */
int pq_trimesh_batch_autotune(pq_trimesh_batch *batch, const pq_ray *rays, float max_toi, pq_ray_hit *hits,
                              void *stream, pq_ray_tuning *chosen, float *chosen_ms);
/* Captures one batched cast (with the batch's current tuning) into a hipGraph;
 * replay it with pq_graph_launch.  stream NULL: an internal capture stream. */
int pq_trimesh_batch_capture(const pq_trimesh_batch *batch, const pq_ray *rays, float max_toi, int solid,
                             pq_ray_hit *hits, void *stream, pq_graph **out);
int pq_graph_launch(pq_graph *graph, void *stream);
int pq_graph_destroy(pq_graph *graph);

/* ---- Shapes and poses ---------------------------------------------------- */

/* parry3d shape types on the path (ShapeType, src/shape/shape.rs:36-94). */
#define PQ_SHAPE_BALL 0u              /* params[0] = radius (src/shape/ball.rs:50-56) */
#define PQ_SHAPE_CUBOID 1u            /* params[0..3] = half extents (src/shape/cuboid.rs:69-79) */
#define PQ_SHAPE_CAPSULE 2u           /* params[0] = half height along y, params[1] = radius (capsule.rs:76-82) */
#define PQ_SHAPE_CONE 3u              /* params[0] = half height along y, params[1] = radius (cone.rs:80-92) */
#define PQ_SHAPE_CONVEX_POLYHEDRON 4u /* vertex range [first_vertex, +num_vertices) (convex_polyhedron.rs:172-185);
                                         params must be 0 (pq_shapes_create rejects others) */
#define PQ_SHAPE_TRIANGLE 5u          /* vertices a, b, c = [first_vertex, +3) (src/shape/triangle.rs) */

typedef struct pq_shape_desc {
    uint32_t type;
    uint32_t first_vertex;
    uint32_t num_vertices;
    uint32_t reserved; /* must be 0 */
    float params[4];
} pq_shape_desc;

/* nalgebra Isometry3<f32> (#[repr(C)]: UnitQuaternion coords [i, j, k, w], then
 * the translation), 28 bytes: a Rust `&[Isometry3<f32>]` can be passed as is. */
typedef struct pq_isometry {
    float rotation[4];
    float translation[3];
} pq_isometry;

typedef struct pq_shapes pq_shapes;

/* Uploads a shape set: `shapes` (num_shapes records) and the ConvexPolyhedron
 * vertex pool (`vertices`, 3*num_vertices floats).  Host pointers, copied. */
int pq_shapes_create(const pq_shape_desc *shapes, uint32_t num_shapes, const float *vertices,
                     uint32_t num_vertices, pq_shapes **out);
int pq_shapes_destroy(pq_shapes *shapes);

/* Shape::compute_aabb(pos) for every shape (src/shape/shape.rs:369 and the
 * per-type aabb functions of src/bounding_volume/aabb_*.rs).  poses: device,
 * num_shapes pq_isometry; aabbs: device, num_shapes x {mins[3], maxs[3]}. */
int pq_shapes_compute_aabbs(const pq_shapes *shapes, const pq_isometry *poses, float *aabbs, void *stream);

/* ---- Broadphase Bvh ------------------------------------------------------- */

typedef struct pq_bvh pq_bvh;

/* Bvh::from_leaves(BvhBuildStrategy::Binned, aabbs) (bvh_tree.rs:1835-1837) built
 * on the host from `aabbs` (host, n x {mins[3], maxs[3]}) and uploaded. */
int pq_bvh_create(const float *aabbs, uint32_t n, pq_bvh **out);
/* A broadphase handle over a tree the caller already has (host, depth-first
 * wide records over leaves 0..num_leaves-1, e.g. decoded with pq_bvh_decode or
 * exported from a Rust Bvh): the Bvh a deserialized parry Bvh is
 * (bvh_tree.rs:1748-1762).  Fails with PQ_ERR_INVALID if the records are not
 * such a tree. */
int pq_bvh_create_from_nodes(const pq_bvh_node_wide *nodes, uint32_t num_nodes, uint32_t num_leaves, pq_bvh **out);
/* Bvh::from_leaves(BvhBuildStrategy::Ploc, aabbs) (bvh_tree.rs:1891-1955 with
 * rebuild_range_ploc, bvh_ploc_build.rs:10-93, and refit's depth-first layout)
 * built on the GPU, from `aabbs` in device memory, into a broadphase handle. */
int pq_bvh_create_ploc(const float *aabbs, uint32_t n, pq_bvh **out, void *stream);
int pq_bvh_destroy(pq_bvh *bvh);
int pq_bvh_info(const pq_bvh *bvh, uint32_t *num_leaves, uint32_t *num_nodes, uint32_t *depth);
/* Copies the device tree to host memory (synchronises the default stream). */
int pq_bvh_export(const pq_bvh *bvh, pq_bvh_node_wide *out_host, uint32_t capacity);

/* One frame's tree update: Bvh::insert_or_update_partially(aabbs[i], i, 0.0) for
 * every leaf i (bvh_insert.rs:209-231) followed by Bvh::refit (bvh_refit.rs:170-320;
 * the depth-first layout is already in place, so only boxes, counts and change
 * flags are recomputed).  aabbs: device, n x {mins[3], maxs[3]}.  Calls that
 * update one Bvh must be ordered (one stream), as parry's `&mut self`. */
int pq_bvh_update_and_refit(pq_bvh *bvh, const float *aabbs, void *stream);

/* Bvh::traverse_bvtt_single_tree::<false> (bvh_traverse_bvtt.rs:19-204): every
 * unordered pair of leaves whose AABBs overlap (closed intervals,
 * BvhNode::intersects bvh_tree.rs:954-957), written as (i, j) with i < j, grouped
 * by i in increasing order.  pairs: device, room for `capacity` pairs (2 u32 each);
 * count: device u64 receiving the total number of pairs (also when it exceeds
 * `capacity`, in which case only the first `capacity` are written).  The set
 * equals parry's callback set; the order of the list is this function's. */
int pq_bvh_self_pairs(pq_bvh *bvh, uint32_t *pairs, uint64_t capacity, uint64_t *count, void *stream);

/* ---- Narrowphase contacts ------------------------------------------------ */

/* parry3d::query::Contact (src/query/contact/contact.rs:71-105) in world space,
 * plus `has_contact`: 1 = Some(contact), 0 = None (no contact within
 * `prediction`), PQ_CONTACT_UNSUPPORTED = the pair's shape types are not
 * handled by this build. */
#define PQ_CONTACT_UNSUPPORTED 0xFFFFFFFFu
typedef struct pq_contact {
    float point1[3];
    float point2[3];
    float normal1[3];
    float normal2[3];
    float dist;
    uint32_t has_contact;
} pq_contact;

/* Kernel mapping of the contact path (no counterpart in parry; results are
 * identical for every setting).  Process-wide.
 *   epa_lanes_per_pair: lanes that run one pair's EPA together in the first EPA
 *       tier: 64 (one wave per pair), 32, 16, 8 or 4;
 *   epa_polytope: the first tier's polytope buffers: 0 = 32 vertices / 96 faces
 *       with the ConvexPolyhedra's vertices staged in LDS, 1 = the same without
 *       staging, 2 = 24 vertices / 64 faces, 3 = 32 vertices / 160 faces,
 *       4 = 48 vertices / 224 faces, all three without staging (less LDS per pair
 *       means more pairs resident; a pair whose polytope outgrows the buffers
 *       moves on to the one-wave-per-pair tiers);
 *   epa_schedule: with fewer than 64 lanes per pair, 1 = a wave takes its next
 *       pairs when all its groups are done, 2 = a group of lanes takes its next
 *       pair as soon as it is done with one, 0 = automatic: 1 for the
 *       TriMesh-shape contact (short EPA runs), 2 everywhere else.
 * The default is 16 lanes per pair with polytope 1 and schedule 0. */
typedef struct pq_contact_tuning {
    uint32_t epa_lanes_per_pair;
    uint32_t epa_polytope;
    uint32_t epa_schedule;
    uint32_t reserved; /* must be 0 */
} pq_contact_tuning;
int pq_contact_set_tuning(const pq_contact_tuning *tuning);
/* Phase timing of the contact path, for measurement (no counterpart in parry).
 * While enabled (process-wide), every contact call records HIP events on its
 * own stream between its phases; pq_contact_phase_ms waits for the last call's
 * events and writes ms[0..4]: the GJK/dispatch kernel, the first EPA tier, the
 * 768-face tier, the 2048-face tier and the global-memory tier.  Calls that
 * overlap on several streams share the events; leave it off under graph capture. */
int pq_contact_set_phase_timing(int enable);
int pq_contact_phase_ms(float *ms, uint32_t n);
int pq_contact_get_tuning(pq_contact_tuning *tuning);

/* Device workspace bytes pq_contact_pairs needs for `num_pairs` pairs. */
size_t pq_contact_workspace_bytes(uint64_t num_pairs);

/* ---- ConvexPolyhedron faces (for contact manifolds) --------------------------- */

/* One face of a ConvexPolyhedron as parry stores it (convex_polyhedron.rs:59-63,
 * 172-185): its unit normal, and the range [first, first + count) of
 * face_vertices / face_edges (vertices_adj_to_face, edges_adj_to_face) that lists
 * its vertices (indices into the polyhedron's own points) and edges in order. */
typedef struct pq_polyhedron_face {
    float normal[3];
    uint32_t first;
    uint32_t count;
} pq_polyhedron_face;

/* Gives the set's ConvexPolyhedra their faces: shape_faces holds 2 uint32 per
 * shape (first face in `faces`, face count; count 0 = no faces, and for
 * non-polyhedra it must be 0).  Only contact manifolds read the faces (a
 * ConvexPolyhedron without faces has no feature map here).  Synchronous; replaces
 * faces given before. */
int pq_shapes_set_polyhedron_faces(pq_shapes *shapes, const pq_polyhedron_face *faces, uint32_t num_faces,
                                   const uint32_t *face_vertices, const uint32_t *face_edges, uint32_t num_entries,
                                   const uint32_t *shape_faces);

/* ---- Contact manifolds between polygonal feature maps ------------------------ */

/* TrackedContact (src/query/contact_manifolds/contact_manifold.rs:16-60): the
 * points in the local frames of shape 1 and shape 2, the signed distance and the
 * PackedFeatureId of each shape's feature (feature_id.rs:446-543; 0 = unknown). */
typedef struct pq_tracked_contact {
    float local_p1[3];
    float local_p2[3];
    float dist;
    uint32_t fid1;
    uint32_t fid2;
} pq_tracked_contact;

/* At most 25 points: 4 + 4 vertices inside the other face, 16 edge crossings, and
 * GJK's own point. */
#define PQ_MANIFOLD_MAX_POINTS 25
/* ContactManifold (contact_manifold.rs:470-500) as contact_manifold_pfm_pfm leaves
 * it.  status: 1 = GJK found closest points (the points are filled); 0 = no
 * contact within the prediction (no points; local_n1 caches GJK's separating
 * direction, as parry's manifold does); PQ_CONTACT_UNSUPPORTED = a shape has no
 * polygonal feature map here. */
typedef struct pq_contact_manifold {
    float local_n1[3];
    float local_n2[3];
    uint32_t num_points;
    uint32_t status;
    pq_tracked_contact points[PQ_MANIFOLD_MAX_POINTS];
} pq_contact_manifold;

/* Batched contact_manifold_pfm_pfm(pos12, pfm1, border1, None, pfm2, border2, None,
 * prediction, manifold) (contact_manifolds_pfm_pfm.rs:44-162) on a fresh manifold
 * (no points, zero cached normal: the first contact of a pair, or every contact
 * whose manifold update fails) for every (i, j) of `pairs`.  Feature maps
 * (shape.rs as_polygonal_feature_map): Cuboid, Cone and Triangle are their own; a
 * Capsule is its segment with the radius as border; a ConvexPolyhedron is its own
 * once pq_shapes_set_polyhedron_faces gave it faces; a Ball has none.  parry's dispatcher sends Cuboid-Cuboid, Capsule-Capsule and
 * Cuboid-Triangle to other routines; this call runs the feature-map routine on
 * whatever pairs it is given.
 *   poses: device, one pq_isometry per shape; pairs: device, num_pairs x (i, j);
 *   out: device, one pq_contact_manifold per pair; workspace: device,
 *   pq_contact_manifolds_workspace_bytes(num_pairs) bytes, or NULL.  Asynchronous
 *   on `stream`. */
int pq_contact_manifolds_pfm_pairs(const pq_shapes *shapes, const pq_isometry *poses, const uint32_t *pairs,
                                   uint64_t num_pairs, float prediction, pq_contact_manifold *out, void *workspace,
                                   void *stream);
size_t pq_contact_manifolds_workspace_bytes(uint64_t num_pairs);

/* Status of a manifold that ContactManifold::try_update_contacts_eps kept: its
 * points (and their feature ids) are last frame's, with local_p1 and dist
 * updated to the new poses, so a caller's per-point data stays with them. */
#define PQ_MANIFOLD_UPDATED 2
/* contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162) on the manifolds
 * of the previous frame, in place: `manifolds` holds one pq_contact_manifold per
 * pair as the last call left it (all zero for a new pair: a fresh manifold).  A
 * manifold that try_update_contacts_eps(pos12, COS_1_DEGREES, 1e-6) keeps
 * (contact_manifold.rs:662-699) is updated and gets PQ_MANIFOLD_UPDATED; every
 * other pair is recomputed as pq_contact_manifolds_pfm_pairs does, with GJK seeded
 * by the cached normal local_n1 (parry's default direction if it is zero).
 * parry's match_contacts then moves a caller's per-point data from the old
 * points to new ones with the same (fid1, fid2); that data is the caller's, so
 * the caller does the matching, from its copy of the old points, for the pairs
 * whose status is not PQ_MANIFOLD_UPDATED.  Workspace:
 * pq_contact_manifolds_workspace_bytes(num_pairs) bytes, or NULL.  Asynchronous on
 * `stream`. */
int pq_contact_manifolds_pfm_update(const pq_shapes *shapes, const pq_isometry *poses, const uint32_t *pairs,
                                    uint64_t num_pairs, float prediction, pq_contact_manifold *manifolds,
                                    void *workspace, void *stream);

/* ---- Linear shape casts ---------------------------------------------------- */

/* ShapeCastOptions (src/query/shape_cast/shape_cast.rs:105-146).  parry's
 * default: max_time_of_impact = FLT_MAX, target_distance = 0,
 * stop_at_penetration = 1, compute_impact_geometry_on_penetration = 1. */
typedef struct pq_shape_cast_options {
    float max_time_of_impact;
    float target_distance;
    uint32_t stop_at_penetration;
    uint32_t compute_impact_geometry_on_penetration;
} pq_shape_cast_options;

/* Option<ShapeCastHit> (shape_cast.rs:36-70): status PQ_CAST_NONE is None,
 * otherwise ShapeCastStatus (shape_cast.rs:12-33).  witness1 and normal1 are
 * in shape 1's local frame, witness2 and normal2 in shape 2's. */
#define PQ_CAST_NONE 0u
#define PQ_CAST_OUT_OF_ITERATIONS 1u
#define PQ_CAST_CONVERGED 2u
#define PQ_CAST_FAILED 3u
#define PQ_CAST_PENETRATING 4u
typedef struct pq_shape_cast_hit {
    float time_of_impact;
    float witness1[3];
    float witness2[3];
    float normal1[3];
    float normal2[3];
    uint32_t status;
} pq_shape_cast_hit;

/* Batched query::cast_shapes(pos1, vel1, g1, pos2, vel2, g2, options)
 * (shape_cast.rs:274-286) with DefaultQueryDispatcher::cast_shapes's routing
 * (default_query_dispatcher.rs:434-540): Ball-Ball in closed form
 * (shape_cast_ball_ball.rs:10-68); every other pair of the shape set by GJK's
 * ray cast on the Minkowski difference (cast_shapes_support_map_support_map,
 * shape_cast_support_map_support_map.rs:11-72; gjk.rs:632-795), with the
 * impact geometry of penetrating pairs from contact_support_map_support_map.
 *   poses: device, one pq_isometry per shape; vels: device, 3 floats (linear
 *   velocity) per shape; pairs: device, num_pairs x (i, j);
 *   out: device, one pq_shape_cast_hit per pair;
 *   iters: device, nullable, GJK ray-cast iterations per pair (0 for Ball-Ball);
 *   workspace: device, pq_cast_shapes_workspace_bytes(num_pairs) bytes, or NULL
 *   to let the library allocate it on `stream`.  Asynchronous on `stream`. */
int pq_cast_shapes_pairs(const pq_shapes *shapes, const pq_isometry *poses, const float *vels, const uint32_t *pairs,
                         uint64_t num_pairs, const pq_shape_cast_options *options, pq_shape_cast_hit *out,
                         uint32_t *iters, void *workspace, void *stream);
/* Device workspace bytes pq_cast_shapes_pairs needs for `num_pairs` pairs. */
size_t pq_cast_shapes_workspace_bytes(uint64_t num_pairs);

/* ---- Contacts between a TriMesh and convex shapes ------------------------ */

/* query::contact(mesh_pose, TriMesh, poses[q], shape q, prediction) for every
 * shape q of a shape set: the composite dispatch (default_query_dispatcher.rs:343-346)
 * to CompositeShapeRef::contact_with_shape (contact_composite_shape_shape.rs:14-42).
 * Shape q's AABB in the mesh's frame, loosened by `prediction`, selects
 * triangles in Bvh::leaves order (bvh_traverse.rs:1-68); each gets the contact
 * dispatch as a Triangle; the smallest dist wins (the first on ties).  Results
 * are in world space; normal1 and point1 belong to the mesh.
 * pq_mesh_contact_create copies the mesh's triangles and the shapes into one
 * device shape table; the mesh and shapes must outlive the handle. */
typedef struct pq_mesh_contact pq_mesh_contact;
int pq_mesh_contact_create(const pq_trimesh *mesh, const pq_shapes *shapes, pq_mesh_contact **out);
int pq_mesh_contact_destroy(pq_mesh_contact *mc);
/* Device workspace for up to `max_candidates` (triangle, shape) candidates per call. */
size_t pq_mesh_contact_workspace_bytes(const pq_mesh_contact *mc, uint64_t max_candidates);
/* Candidates are numbered shape by shape; a shape with a candidate numbered
 * `max_candidates` or above gets has_contact = PQ_CONTACT_TRUNCATED (the shapes
 * before it are complete); *num_candidates (device, nullable) receives the
 * number the call needed. */
#define PQ_CONTACT_TRUNCATED 0xFFFFFFFEu
/*   mesh_pose: device, one pq_isometry; poses: device, one per shape;
 *   out: device, one pq_contact per shape; triangles: device, nullable, the
 *   index of the triangle in contact (PQ_NO_HIT for None).  Asynchronous on
 *   `stream`; no allocation. */
int pq_mesh_contact_shapes(const pq_mesh_contact *mc, const float *mesh_pose, const float *poses, float prediction,
                           uint64_t max_candidates, pq_contact *out, uint32_t *triangles, uint64_t *num_candidates,
                           void *workspace, void *stream);

/* Batched query::contact(pos1, g1, pos2, g2, prediction)
 * (src/query/contact/contact_shape_shape.rs:123-138) over shape pairs, with
 * DefaultQueryDispatcher::contact's routing (default_query_dispatcher.rs:302-356):
 * Ball-Ball closed form (contact_ball_ball.rs:9-42); Ball with Cuboid, Capsule,
 * Cone or ConvexPolyhedron by projecting the ball centre
 * (contact_ball_convex_polyhedron.rs:12-63) -- Ball with Triangle too, by the
 * triangle's closed-form projection; every other pair by GJK/EPA
 * (contact_support_map_support_map.rs:10-77).
 *   shapes:    the shape set; poses: device, one pq_isometry per shape;
 *   pairs:     device, num_pairs x (i, j) shape indices;
 *   init_dirs: device, nullable; num_pairs x 3 floats: the GJK seed direction
 *              (unit, in shape i's local frame) passed as `init_dir` to
 *              contact_support_map_support_map_with_params
 *              (contact_support_map_support_map.rs:40-60); NULL = parry's
 *              default seed normalize(pos12.translation), else +x.  Ignored
 *              for pairs with a Ball (parry takes no seed there);
 *   out:       device, num_pairs pq_contact;
 *   iters:     device, nullable, 2 u32 per pair: GJK iterations and EPA
 *              iterations (0 when EPA did not run);
 *   workspace: device, nullable, pq_contact_workspace_bytes(num_pairs) bytes;
 *              NULL = a library-owned per-stream buffer, grown on demand (the
 *              growth allocates: give a workspace under hipGraph capture).
 * Asynchronous on `stream`. */
int pq_contact_pairs(const pq_shapes *shapes, const pq_isometry *poses, const uint32_t *pairs, uint64_t num_pairs,
                     float prediction, const float *init_dirs, pq_contact *out, uint32_t *iters, void *workspace,
                     void *stream);

/* Same as pq_contact_pairs for the first min(capacity, *num_pairs) pairs, where
 * num_pairs is a device u64 (e.g. the count written by pq_bvh_self_pairs), so
 * a broadphase + narrowphase frame runs without a host round trip.  Records
 * beyond the count are not written.  workspace: pq_contact_workspace_bytes(capacity). */
int pq_contact_pairs_device_count(const pq_shapes *shapes, const pq_isometry *poses, const uint32_t *pairs,
                                  uint64_t capacity, const uint64_t *num_pairs, float prediction,
                                  const float *init_dirs, pq_contact *out, uint32_t *iters, void *workspace,
                                  void *stream);

/* pq_contact_pairs without any device allocation.  No counterpart in parry: it
 * is for callers that cannot allocate (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc).  The workspace is a fixed 128 MiB
 * device array inside the library, one per GPU, held until the process ends;
 * num_pairs above PQ_CONTACT_NOALLOC_MAX_PAIRS is rejected with PQ_ERR_INVALID.
 * Results are those of pq_contact_pairs.  Every call on one GPU uses that GPU's
 * array: calls that overlap on one GPU (several streams or host threads) race.
 * Asynchronous on `stream`. */
#define PQ_CONTACT_NOALLOC_MAX_PAIRS 262144u
int pq_contact_pairs_noalloc(const pq_shapes *shapes, const pq_isometry *poses, const uint32_t *pairs,
                             uint64_t num_pairs, float prediction, const float *init_dirs, pq_contact *out,
                             uint32_t *iters, void *stream);

/* ---- Learned GJK warm start (no counterpart in parry) ------------------------
 * A 44-128-128-128-3 MLP (ReLU, bf16 weights on MFMA with f32 accumulation)
 * predicts the GJK seed direction that pq_contact_pairs takes as `init_dirs`
 * (the init_dir slot of contact_support_map_support_map_with_params,
 * contact_support_map_support_map.rs:52-60).  The seed only starts the simplex:
 * GJK still stops at parry's tolerance.  A zero seed means parry's default seed,
 * and a seeded GJK that reaches the 100-iteration cap is re-run with the
 * default seed.
 *   descriptors: 16 floats per shape (pq_shape_descriptors);
 *   features:    PQ_MLP_FEATURE_STRIDE floats per pair: the rotation matrix
 *                (row-major) and translation of pos_i^-1 * pos_j, then the
 *                descriptors of shapes i and j, then 4 zeros;
 *   parameters:  PQ_MLP_NUM_PARAMS floats: W1 (128x44, row-major), b1, W2
 *                (128x128), b2, W3, b3, W4 (3x128), b4.
 * All buffers except `params` / `epoch_losses` are device pointers. */
#define PQ_MLP_NUM_PARAMS 39171u
#define PQ_MLP_FEATURES 44u
#define PQ_MLP_FEATURE_STRIDE 48u
#define PQ_SHAPE_DESCRIPTOR_FLOATS 16u
typedef struct pq_mlp pq_mlp;
typedef struct pq_mlp_train_config {
    uint32_t epochs;
    uint32_t batch;    /* multiple of 32, at most 65536 */
    float lr;          /* Adam learning rate, cosine-decayed to lr_final */
    float lr_final;
    float beta1, beta2, eps;
    uint32_t reserved; /* must be 0 */
    uint64_t shuffle_seed;
} pq_mlp_train_config;

/* params: host, PQ_MLP_NUM_PARAMS floats, or NULL for a He initialisation from `seed`. */
int pq_mlp_create(const float *params, uint64_t seed, pq_mlp **out);
int pq_mlp_destroy(pq_mlp *net);
int pq_mlp_get_params(const pq_mlp *net, float *params);
int pq_mlp_set_params(pq_mlp *net, const float *params); /* also resets the Adam state */
int pq_shape_descriptors(const pq_shapes *shapes, float *descriptors, void *stream);
int pq_pair_features(const pq_isometry *poses, const uint32_t *pairs, uint64_t num_pairs, const float *descriptors,
                     float *features, void *stream);
/* Raw network outputs (n x 3) for feature rows. */
int pq_mlp_forward(const pq_mlp *net, const float *features, uint64_t n, float *out, void *stream);
/* Features + network + normalisation fused: unit seeds (num_pairs x 3; zero where
 * the output is degenerate), ready for pq_contact_pairs' init_dirs. */
int pq_contact_seeds(const pq_mlp *net, const pq_isometry *poses, const uint32_t *pairs, uint64_t num_pairs,
                     const float *descriptors, float *seeds, void *stream);
/* Loss gradient over feature rows `rows[0..count)` with unit labels (3 per row):
 * grads (PQ_MLP_NUM_PARAMS floats), loss_sum (one double: the summed 1 - cos). */
int pq_mlp_gradients(pq_mlp *net, const float *features, const float *labels, const uint32_t *rows, uint32_t count,
                     float *grads, double *loss_sum, void *stream);
/* Adam training on n rows; epoch_losses (host, nullable): mean loss per epoch.
 * Synchronises `stream`. */
int pq_mlp_train(pq_mlp *net, const float *features, const float *labels, uint64_t n, const pq_mlp_train_config *cfg,
                 double *epoch_losses, void *stream);
/* normal1 of each contact rotated into shape i's frame (zero where has_contact != 1):
 * the plain GJK's direction, the label the network is trained on. */
int pq_contact_local_normals(const pq_contact *contacts, const pq_isometry *poses, const uint32_t *pairs,
                             uint64_t num_pairs, float *normals, void *stream);

/* Same as pq_contact_seeds for the first min(capacity, *num_pairs) pairs, the
 * count being a device u64 (e.g. written by pq_bvh_self_pairs). */
int pq_contact_seeds_device_count(const pq_mlp *net, const pq_isometry *poses, const uint32_t *pairs,
                                  uint64_t capacity, const uint64_t *num_pairs, const float *descriptors, float *seeds,
                                  void *stream);

/* ---- Simulation frames (no counterpart in parry's query path) ----------------
 * Rigid motion: t += v dt; q = normalize(dq * q), dq the rotation by |w| dt about
 * the world-frame angular velocity w.  linvel, angvel: device, 3 floats per shape. */
int pq_integrate_poses(pq_isometry *poses, const float *linvel, const float *angvel, uint32_t n, float dt,
                       void *stream);
/* Per-pair warm-start cache: (i, j) -> last contact normal1 in shape i's frame,
 * the next frame's GJK seed (parry's manifolds reuse their previous normal the
 * same way, contact_manifolds_pfm_pfm.rs:67).  A device buffer of
 * pq_warm_cache_bytes(capacity) bytes, capacity a power of two well above the
 * pair count; cleared by a kernel (capturable in a graph). */
size_t pq_warm_cache_bytes(uint32_t capacity);
/* Zeroes `bytes` (a multiple of 4) of device memory with a kernel (no memset node
 * when captured). */
int pq_device_zero(void *ptr, size_t bytes, void *stream);
int pq_warm_cache_clear(void *cache, uint32_t capacity, void *stream);
/* Stores normal1 of every pair with has_contact == 1 (first min(pair_capacity, *num_pairs) pairs). */
int pq_warm_cache_store(void *cache, uint32_t capacity, const uint32_t *pairs, const uint64_t *num_pairs,
                        uint64_t pair_capacity, const pq_contact *contacts, const pq_isometry *poses, void *stream);
/* Overwrites seeds[k] with the cached normal of pair k where there is one;
 * hits (device u32, nullable) counts them. */
int pq_warm_cache_apply(const void *cache, uint32_t capacity, const uint32_t *pairs, const uint64_t *num_pairs,
                        uint64_t pair_capacity, float *seeds, uint32_t *hits, void *stream);
/* Records the library calls made on `stream` between the two calls into a
 * hipGraph (replay with pq_graph_launch).  The calls must not allocate: pass
 * workspaces, and capacities large enough. */
int pq_stream_begin_capture(void *stream);
int pq_stream_end_capture(void *stream, pq_graph **out);

#ifdef __cplusplus
}
#endif

#endif /* PARRY_MI355X_H */
