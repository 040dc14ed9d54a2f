"""Query types mirroring parry3d ``src/query`` (ray casts; contacts are added beside them)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

Vec3 = Tuple[float, float, float]

# parry3d::query::Ray, #[repr(C)] f32 (src/query/ray/ray.rs:75-88).
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3)])
# pq_ray_hit (include/parry_mi355x.h): RayIntersection with the TriMesh feature index.
RAY_HIT_DTYPE = np.dtype([("time_of_impact", "<f4"), ("normal", "<f4", 3), ("feature", "<u4")])
assert RAY_DTYPE.itemsize == 24 and RAY_HIT_DTYPE.itemsize == 20


@dataclass(frozen=True)
class Ray:
    """parry3d ``Ray`` (ray.rs:67-88): origin and (not necessarily unit) direction."""

    origin: Vec3
    dir: Vec3

    def point_at(self, t: float) -> Vec3:
        return tuple(o + d * t for o, d in zip(self.origin, self.dir))


@dataclass(frozen=True)
class FeatureId:
    """parry3d ``FeatureId`` (src/shape/feature_id.rs:117): kind in {Vertex, Edge, Face, Unknown}."""

    kind: str
    index: int = 0

    @staticmethod
    def face(i: int) -> "FeatureId":
        return FeatureId("Face", i)


@dataclass(frozen=True)
class RayIntersection:
    """parry3d ``RayIntersection`` (ray.rs:286-315)."""

    time_of_impact: float
    normal: Vec3
    feature: FeatureId
