"""parry_amd -- MI355X (gfx950) batched collision queries behind parry3d's API.

The compute path is the HIP library ``libparry_mi355x.so`` (C ABI in
include/parry_mi355x.h); this package is its Python host mirror.
"""
from .query import (
    Ray,
    RayIntersection,
    FeatureId,
    RAY_DTYPE,
    RAY_HIT_DTYPE,
    Contact,
    CONTACT_DTYPE,
    CONTACT_UNSUPPORTED,
    contact_pairs,
    contact_workspace_bytes,
    MeshContact,
    PQ_CONTACT_TRUNCATED,
    PQ_CAST_CONVERGED,
    PQ_CAST_NONE,
    PQ_CAST_PENETRATING,
    cast_shapes_pairs,
    cast_shapes_workspace_bytes,
)
from .shape import (
    TriMesh,
    TriMeshBatch,
    Graph,
    TriMeshBuilderError,
    BVH_NODE_WIDE_DTYPE,
    split_hits,
    split_projections,
    ShapeSet,
    SHAPE_DESC_DTYPE,
    SHAPE_BALL,
    SHAPE_CUBOID,
    SHAPE_CAPSULE,
    SHAPE_CONE,
    SHAPE_CONVEX_POLYHEDRON,
    SHAPE_TRIANGLE,
)
from .partitioning import Bvh, ploc_nodes
from . import scenes

__all__ = [
    "Ray",
    "RayIntersection",
    "FeatureId",
    "RAY_DTYPE",
    "RAY_HIT_DTYPE",
    "TriMesh",
    "TriMeshBatch",
    "Graph",
    "TriMeshBuilderError",
    "BVH_NODE_WIDE_DTYPE",
    "split_hits",
    "split_projections",
    "ShapeSet",
    "SHAPE_DESC_DTYPE",
    "Bvh",
    "scenes",
    "Contact",
    "CONTACT_DTYPE",
    "contact_pairs",
    "contact_workspace_bytes",
    "ploc_nodes",
    "MeshContact",
    "PQ_CONTACT_TRUNCATED",
    "PQ_CAST_CONVERGED",
    "PQ_CAST_NONE",
    "PQ_CAST_PENETRATING",
    "cast_shapes_pairs",
    "cast_shapes_workspace_bytes",
]
