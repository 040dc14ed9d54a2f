"""ctypes binding of ``libparry_mi355x.so`` (the C ABI in include/parry_mi355x.h).

The library is required: importing a function from here when the shared
object is missing raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparry_mi355x.so")

PQ_OK = 0
PQ_ERR_INVALID = 1
PQ_ERR_UNSUPPORTED = 2
PQ_ERR_OOM = 3
PQ_ERR_HIP = 4
PQ_ERR_CAPACITY = 5
PQ_NO_HIT = 0xFFFFFFFF

_c = ctypes
_vp = _c.c_void_p
_u32 = _c.c_uint32
_u64 = _c.c_uint64
_f32 = _c.c_float
_int = _c.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "pq_last_error": (_c.c_size_t, [_c.c_char_p, _c.c_size_t]),
    "pq_device_malloc": (_int, [_c.POINTER(_vp), _c.c_size_t]),
    "pq_device_free": (_int, [_vp]),
    "pq_memcpy_htod": (_int, [_vp, _vp, _c.c_size_t, _vp]),
    "pq_memcpy_dtoh": (_int, [_vp, _vp, _c.c_size_t, _vp]),
    "pq_stream_synchronize": (_int, [_vp]),
    "pq_bvh_from_leaves": (_int, [_vp, _u32, _vp, _u32, _c.POINTER(_u32), _c.POINTER(_u32)]),
    "pq_trimesh_create": (_int, [_vp, _u32, _vp, _u32, _vp, _u32, _c.POINTER(_vp)]),
    "pq_trimesh_destroy": (_int, [_vp]),
    "pq_trimesh_info": (_int, [_vp, _c.POINTER(_u32), _c.POINTER(_u32), _c.POINTER(_u32)]),
    "pq_trimesh_set_tuning": (_int, [_vp, _vp]),
    "pq_trimesh_get_tuning": (_int, [_vp, _vp]),
    "pq_trimesh_autotune": (_int, [_vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp]),
    "pq_trimesh_export_bvh": (_int, [_vp, _vp, _u32]),
    "pq_trimesh_cast_local_rays": (_int, [_vp, _vp, _u64, _f32, _int, _vp, _vp]),
    "pq_trimesh_cast_local_rays_counted": (_int, [_vp, _vp, _u64, _f32, _int, _vp, _vp, _vp]),
    "pq_trimesh_cast_local_rays_host": (_int, [_vp, _vp, _u64, _f32, _int, _vp]),
    "pq_trimesh_cast_local_rays_with_culling": (_int, [_vp, _vp, _u64, _f32, _u32, _vp, _vp]),
    "pq_trimesh_project_local_points": (_int, [_vp, _vp, _u64, _int, _vp, _vp]),
    "pq_mesh_contact_create": (_int, [_vp, _vp, _c.POINTER(_vp)]),
    "pq_mesh_contact_destroy": (_int, [_vp]),
    "pq_mesh_contact_workspace_bytes": (_c.c_size_t, [_vp, _u64]),
    "pq_mesh_contact_shapes": (_int, [_vp, _vp, _vp, _f32, _u64, _vp, _vp, _vp, _vp, _vp]),
    "pq_trimesh_project_local_points_counted": (_int, [_vp, _vp, _u64, _int, _vp, _vp, _vp]),
    "pq_trimesh_cast_local_rays_with_culling_counted": (_int, [_vp, _vp, _u64, _f32, _u32, _vp, _vp, _vp]),
    "pq_trimesh_batch_create": (_int, [_vp, _u32, _vp, _c.POINTER(_vp)]),
    "pq_trimesh_batch_destroy": (_int, [_vp]),
    "pq_trimesh_batch_set_tuning": (_int, [_vp, _vp]),
    "pq_trimesh_batch_get_tuning": (_int, [_vp, _vp]),
    "pq_trimesh_batch_cast_local_rays": (_int, [_vp, _vp, _f32, _int, _vp, _vp]),
    "pq_trimesh_batch_cast_local_rays_counted": (_int, [_vp, _vp, _f32, _int, _vp, _vp, _vp]),
    "pq_trimesh_batch_autotune": (_int, [_vp, _vp, _f32, _vp, _vp, _vp, _c.POINTER(_f32)]),
    "pq_trimesh_batch_capture": (_int, [_vp, _vp, _f32, _int, _vp, _vp, _c.POINTER(_vp)]),
    "pq_graph_launch": (_int, [_vp, _vp]),
    "pq_graph_destroy": (_int, [_vp]),
    "pq_shapes_create": (_int, [_vp, _u32, _vp, _u32, _c.POINTER(_vp)]),
    "pq_shapes_destroy": (_int, [_vp]),
    "pq_shapes_set_polyhedron_faces": (_int, [_vp, _vp, _u32, _vp, _vp, _u32, _vp]),
    "pq_shapes_compute_aabbs": (_int, [_vp, _vp, _vp, _vp]),
    "pq_bvh_create": (_int, [_vp, _u32, _c.POINTER(_vp)]),
    "pq_bvh_create_from_nodes": (_int, [_vp, _u32, _u32, _c.POINTER(_vp)]),
    "pq_bvh_encode": (_int, [_vp, _u32, _vp, _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "pq_bvh_decode": (_int, [_vp, _c.c_size_t, _vp, _u32, _c.POINTER(_u32)]),
    "pq_trimesh_encode": (_int, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "pq_trimesh_decode": (_int, [_vp, _c.c_size_t, _vp, _u32, _c.POINTER(_u32), _vp, _u32, _c.POINTER(_u32), _vp, _u32,
                                 _c.POINTER(_u32)]),
    "pq_bvh_destroy": (_int, [_vp]),
    "pq_bvh_create_ploc": (_int, [_vp, _u32, _c.POINTER(_vp), _vp]),
    "pq_bvh_ploc_workspace_bytes": (_c.c_size_t, [_u32]),
    "pq_bvh_from_leaves_ploc": (_int, [_vp, _u32, _vp, _u32, _c.POINTER(_u32), _c.POINTER(_u32), _vp, _vp]),
    "pq_bvh_info": (_int, [_vp, _c.POINTER(_u32), _c.POINTER(_u32), _c.POINTER(_u32)]),
    "pq_bvh_export": (_int, [_vp, _vp, _u32]),
    "pq_bvh_update_and_refit": (_int, [_vp, _vp, _vp]),
    "pq_bvh_self_pairs": (_int, [_vp, _vp, _u64, _vp, _vp]),
    "pq_contact_workspace_bytes": (_c.c_size_t, [_u64]),
    "pq_contact_manifolds_workspace_bytes": (_c.c_size_t, [_u64]),
    "pq_contact_manifolds_pfm_pairs": (_int, [_vp, _vp, _vp, _u64, _f32, _vp, _vp, _vp]),
    "pq_contact_manifolds_pfm_update": (_int, [_vp, _vp, _vp, _u64, _f32, _vp, _vp, _vp]),
    "pq_cast_shapes_workspace_bytes": (_c.c_size_t, [_u64]),
    "pq_cast_shapes_pairs": (_int, [_vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    "pq_contact_set_tuning": (_int, [_vp]),
    "pq_contact_get_tuning": (_int, [_vp]),
    "pq_contact_set_phase_timing": (_int, [_int]),
    "pq_contact_phase_ms": (_int, [_vp, _u32]),
    "pq_contact_pairs": (_int, [_vp, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp, _vp]),
    "pq_contact_pairs_device_count": (_int, [_vp, _vp, _vp, _u64, _vp, _f32, _vp, _vp, _vp, _vp, _vp]),
    "pq_contact_pairs_noalloc": (_int, [_vp, _vp, _vp, _u64, _f32, _vp, _vp, _vp, _vp]),
    "pq_mlp_create": (_int, [_vp, _u64, _c.POINTER(_vp)]),
    "pq_mlp_destroy": (_int, [_vp]),
    "pq_mlp_get_params": (_int, [_vp, _vp]),
    "pq_mlp_set_params": (_int, [_vp, _vp]),
    "pq_shape_descriptors": (_int, [_vp, _vp, _vp]),
    "pq_pair_features": (_int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "pq_mlp_forward": (_int, [_vp, _vp, _u64, _vp, _vp]),
    "pq_contact_seeds": (_int, [_vp, _vp, _vp, _u64, _vp, _vp, _vp]),
    "pq_mlp_gradients": (_int, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "pq_mlp_train": (_int, [_vp, _vp, _vp, _u64, _vp, _vp, _vp]),
    "pq_contact_local_normals": (_int, [_vp, _vp, _vp, _u64, _vp, _vp]),
    "pq_contact_seeds_device_count": (_int, [_vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "pq_integrate_poses": (_int, [_vp, _vp, _vp, _u32, _f32, _vp]),
    "pq_warm_cache_bytes": (_c.c_size_t, [_u32]),
    "pq_device_zero": (_int, [_vp, _c.c_size_t, _vp]),
    "pq_warm_cache_clear": (_int, [_vp, _u32, _vp]),
    "pq_warm_cache_store": (_int, [_vp, _u32, _vp, _vp, _u64, _vp, _vp, _vp]),
    "pq_warm_cache_apply": (_int, [_vp, _u32, _vp, _vp, _u64, _vp, _vp, _vp]),
    "pq_stream_begin_capture": (_int, [_vp]),
    "pq_stream_end_capture": (_int, [_vp, _c.POINTER(_vp)]),
}

_lock = threading.Lock()
_lib = None


PQ_RAY_MAPPING_STATIC = 0
PQ_RAY_MAPPING_PERSISTENT = 1
PQ_RAY_MAPPING_WAVE = 2
PQ_RAY_MAPPING_BINNED = 3


class RayTuning(ctypes.Structure):
    """pq_ray_tuning (include/parry_mi355x.h)."""

    _fields_ = [("mapping", _u32), ("chunk", _u32), ("blocks_per_cu", _u32), ("block_size", _u32),
                ("stack_regs", _u32), ("reserved", _u32)]


RAY_MAPPINGS = {"static": 0, "persistent": 1, "wave": 2, "binned": 3}


def ray_tuning(mapping: str = "persistent", chunk: int = 0, blocks_per_cu: int = 0, block_size: int = 256,
               stack_regs: int = 0) -> RayTuning:
    """A pq_ray_tuning from names: mapping "static", "persistent" or "wave"."""
    return RayTuning(RAY_MAPPINGS[mapping], chunk, blocks_per_cu, block_size, stack_regs, 0)


def ray_tuning_dict(t: RayTuning) -> dict:
    names = {v: k for k, v in RAY_MAPPINGS.items()}
    return {"mapping": names[int(t.mapping)], "chunk": int(t.chunk), "blocks_per_cu": int(t.blocks_per_cu),
            "block_size": int(t.block_size) or 256, "stack_regs": int(t.stack_regs)}


class ContactTuning(ctypes.Structure):
    """pq_contact_tuning (include/parry_mi355x.h)."""

    _fields_ = [("epa_lanes_per_pair", _u32), ("epa_polytope", _u32), ("epa_schedule", _u32), ("reserved", _u32)]


class ShapeCastOptions(ctypes.Structure):
    """pq_shape_cast_options (include/parry_mi355x.h) = parry's ShapeCastOptions."""

    _fields_ = [("max_time_of_impact", _f32), ("target_distance", _f32), ("stop_at_penetration", _u32),
                ("compute_impact_geometry_on_penetration", _u32)]


def set_contact_tuning(epa_lanes_per_pair: int, epa_polytope: int = 1, epa_schedule: int = 0) -> None:
    check(lib().pq_contact_set_tuning(ctypes.byref(ContactTuning(epa_lanes_per_pair, epa_polytope, epa_schedule))))


CONTACT_PHASES = ("gjk", "epa_first_tier", "epa_768_faces", "epa_2048_faces", "epa_global")


def set_contact_phase_timing(enable: bool) -> None:
    """Records HIP events between the contact path's phases on each call's stream."""
    check(lib().pq_contact_set_phase_timing(int(bool(enable))))


def contact_phase_ms() -> dict:
    """Milliseconds of each phase of the last contact call (waits for it)."""
    ms = (ctypes.c_float * len(CONTACT_PHASES))()
    check(lib().pq_contact_phase_ms(ms, len(CONTACT_PHASES)))
    return {k: float(v) for k, v in zip(CONTACT_PHASES, ms)}


def get_contact_tuning() -> tuple:
    t = ContactTuning()
    check(lib().pq_contact_get_tuning(ctypes.byref(t)))
    return int(t.epa_lanes_per_pair), int(t.epa_polytope), int(t.epa_schedule)


class MlpTrainConfig(_c.Structure):
    """pq_mlp_train_config."""

    _fields_ = [("epochs", _u32), ("batch", _u32), ("lr", _f32), ("lr_final", _f32), ("beta1", _f32),
                ("beta2", _f32), ("eps", _f32), ("reserved", _u32), ("shuffle_seed", _u64)]


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pq error {code}] {msg}")
        self.code = code


class Unsupported(NativeError):
    """parry's ``Unsupported`` (src/query/error.rs:146-147)."""


def lib() -> ctypes.CDLL:
    """Loads the native library once; raises if it has not been built.

    torch (when installed) is imported first so that the process has a single
    HIP runtime: torch's bundled libamdhip64.so.7 then also serves this
    library, whose dependency has the same soname, and device pointers and
    streams can be shared with torch tensors.
    """
    global _lib
    with _lock:
        if _lib is None:
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m parry_amd.build` "
                    "(the HIP library is required; there is no fallback)"
                )
            l = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(l, name)
                f.restype = res
                f.argtypes = args
            _lib = l
    return _lib


def last_error() -> str:
    l = lib()
    n = l.pq_last_error(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    l.pq_last_error(buf, n + 1)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != PQ_OK:
        msg = last_error()
        if rc == PQ_ERR_UNSUPPORTED:
            raise Unsupported(rc, msg)
        raise NativeError(rc, msg)
