"""Python binding of the MI355X path tracer's C ABI (include/hpt.h).

`HalaRenderer` mirrors the renderer interface the reference host drives
(`hala_renderer::rt_renderer::HalaRenderer`, called from
src/bin/hala-pathtracer.rs:128-285): the same method names, argument meaning
and validation errors.  Scenes are plain numpy arrays in the std140 layouts
of shaders/src/hala-pathtracer/inc/types.glsl.

The library is the HIP build under hala-pathtracer_amd/lib/libhpt.so; there is
no CPU fallback -- loading fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhpt.so")

# ----------------------------------------------------------------- constants
MAX_CAMERAS = 8
MAX_LIGHTS = 16
INVALID_INDEX = 0xFFFFFFFF
WORLD_SIZE = 10000.0

CAMERA_PERSPECTIVE, CAMERA_ORTHOGRAPHIC = 0, 1
LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_SPOT, LIGHT_QUAD, LIGHT_SPHERE = range(5)
MEDIUM_NONE, MEDIUM_ABSORB, MEDIUM_SCATTER, MEDIUM_EMISSIVE = range(4)
MATERIAL_DIFFUSE, MATERIAL_DISNEY = 0, 1
FEATURE_MEDIUM, FEATURE_VOL_MIS, FEATURE_TRANSPARENT = 1, 2, 4
IMAGE_FINAL_RGBA8, IMAGE_ACCUM, IMAGE_ALBEDO, IMAGE_NORMAL = range(4)
HIT_NONE, HIT_TRIANGLE, HIT_LIGHT = range(3)
STAGE_CAMERA, STAGE_EXTEND, STAGE_SHADE, STAGE_SHADOW, STAGE_ACCUMULATE, STAGE_SORT, STAGE_WALK, STAGE_TAIL = range(8)
PROFILE_TIMING, PROFILE_COUNT_VISITS = 1, 2
TAIL_FROM_NONE = 0xFFFFFFFF
TRAVERSAL_BLOCKS = (64, 128, 256)  # threads per block of the traversal kernels; 128 is the default
LAUNCH_EAGER, LAUNCH_GRAPH = 0, 1
SHADE_SORT_OFF = 0xFFFFFFFF
BVH_HOST, BVH_GPU = 0, 1

STATUS_NAMES = {
    0: "HPT_OK",
    1: "HPT_ERROR_INVALID_ARGUMENT",
    2: "HPT_ERROR_OUT_OF_MEMORY",
    3: "HPT_ERROR_DEVICE",
    4: "HPT_ERROR_STATE",
    5: "HPT_ERROR_UNSUPPORTED",
}

# ------------------------------------------------------- std140 numpy dtypes
CAMERA_DTYPE = np.dtype({
    "names": ["position", "right", "up", "forward", "yfov", "focal_distance_or_xmag", "aperture_or_ymag", "type"],
    "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", "<f4", "<f4", "<u4"],
    "offsets": [0, 16, 32, 48, 60, 64, 68, 72],
    "itemsize": 80,
})
LIGHT_DTYPE = np.dtype({
    "names": ["intensity", "position", "u", "v", "radius", "area", "type"],
    "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", "<f4", "<i4"],
    "offsets": [0, 16, 32, 48, 60, 64, 68],
    "itemsize": 80,
})
MATERIAL_DTYPE = np.dtype({
    "names": ["medium_color", "medium_density", "medium_anisotropy", "medium_type",
              "base_color", "opacity", "emission", "anisotropic", "metallic", "roughness", "subsurface",
              "specular_tint", "sheen", "sheen_tint", "clearcoat", "clearcoat_roughness", "clearcoat_tint",
              "specular_transmission", "ior", "ax", "ay", "base_color_map_index", "normal_map_index",
              "metallic_roughness_map_index", "emission_map_index", "type"],
    "formats": [("<f4", 3), "<f4", "<f4", "<u4", ("<f4", 3), "<f4", ("<f4", 3), "<f4", "<f4", "<f4", "<f4",
                "<f4", "<f4", "<f4", "<f4", "<f4", ("<f4", 3), "<f4", "<f4", "<f4", "<f4", "<u4", "<u4",
                "<u4", "<u4", "<u4"],
    "offsets": [0, 12, 16, 20, 32, 44, 48, 60, 64, 68, 72, 76, 80, 84, 88, 92, 96, 108, 112, 116, 120, 124,
                128, 132, 136, 140],
    "itemsize": 144,
})
RAY_DTYPE = np.dtype({"names": ["origin", "tmin", "direction", "tmax"],
                      "formats": [("<f4", 3), "<f4", ("<f4", 3), "<f4"], "offsets": [0, 12, 16, 28],
                      "itemsize": 32})
HIT_DTYPE = np.dtype({"names": ["t", "u", "v", "kind", "instance", "primitive"],
                      "formats": ["<f4", "<f4", "<f4", "<u4", "<u4", "<u4"], "itemsize": 24})
VERTEX_FLOATS = 11  # position, normal, tangent, tex_coord (types.glsl:86-98)


def new_materials(n: int) -> np.ndarray:
    """Materials with the reference's neutral defaults and no texture maps."""
    m = np.zeros(n, MATERIAL_DTYPE)
    m["opacity"] = 1.0
    m["ior"] = 1.5
    m["medium_anisotropy"] = 0.0
    for k in ("base_color_map_index", "normal_map_index", "metallic_roughness_map_index", "emission_map_index"):
        m[k] = INVALID_INDEX
    m["clearcoat_tint"] = 1.0
    return m


# ------------------------------------------------------------- ctypes ABI
class _Vertex(C.Structure):
    _fields_ = [("data", C.c_float * 11)]


class _PrimitiveDesc(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("material_index", C.c_uint32), ("vertex_count", C.c_uint32),
                ("index_count", C.c_uint32), ("_pad", C.c_uint32), ("vertices", C.c_void_p),
                ("indices", C.c_void_p)]


class _TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba8", C.c_void_p)]


def texture_descs(textures):
    """(hpt_texture_desc array, keep-alive list) for RGBA8 (H, W, 4) arrays."""
    keep = [np.ascontiguousarray(t, np.uint8) for t in textures]
    for t in keep:
        if t.ndim != 3 or t.shape[2] != 4:
            raise ValueError("textures are (height, width, 4) RGBA8 arrays")
    arr = (_TextureDesc * max(1, len(keep)))()
    for i, t in enumerate(keep):
        arr[i] = _TextureDesc(t.shape[1], t.shape[0], t.ctypes.data)
    return arr, keep


class _SceneDesc(C.Structure):
    _fields_ = [("cameras", C.c_void_p), ("camera_count", C.c_uint32), ("light_count", C.c_uint32),
                ("lights", C.c_void_p), ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("primitive_count", C.c_uint32), ("primitives", C.c_void_p)]


class _CreateInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("enable_tonemap", C.c_uint32), ("enable_aces", C.c_uint32),
                ("use_simple_aces", C.c_uint32), ("features", C.c_uint32), ("max_samples", C.c_uint64),
                ("exposure_value", C.c_float), ("device", C.c_int32)]


class _BvhInfo(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("triangle_count", C.c_uint32), ("max_depth", C.c_uint32),
                ("_pad", C.c_uint32)]


class _StageStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("total_ms", C.c_double), ("rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("queries", C.c_uint64)]


class _ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("mode", C.c_uint32), ("budget_bytes", C.c_uint64),
                ("max_frames", C.c_uint32), ("keep_images", C.c_uint32), ("step_paths", C.c_uint64),
                ("launch_mode", C.c_uint32), ("max_running", C.c_uint32), ("dry_run_ns_per_path", C.c_uint64),
                ("reuse_handles", C.c_uint32), ("_reserved", C.c_uint32)]


class _ServeSceneDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("features", C.c_uint32), ("step_launch_mode", C.c_uint32),
                ("step_shade_sort", C.c_uint32), ("step_tail_from", C.c_uint32),
                ("step_traversal_block", C.c_uint32), ("dry_run_fail", C.c_uint32)]


class _JobDesc(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("camera_index", C.c_uint32),
                ("enable_tonemap", C.c_uint32), ("enable_aces", C.c_uint32), ("use_simple_aces", C.c_uint32),
                ("exposure_value", C.c_float), ("frames_in_flight", C.c_uint32), ("launch_mode", C.c_uint32),
                ("shade_sort", C.c_uint32), ("tail_from", C.c_uint32), ("traversal_block", C.c_uint32),
                ("denoise", C.c_uint32)]


class _JobResult(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("status", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("frames", C.c_uint32), ("has_image", C.c_uint32), ("_pad", C.c_uint32),
                ("footprint_bytes", C.c_uint64), ("submitted_s", C.c_double), ("admitted_s", C.c_double),
                ("finished_s", C.c_double), ("error", C.c_char * 256)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("finished", C.c_uint64), ("failed", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64), ("wavefront_steps", C.c_uint64),
                ("max_jobs_per_step", C.c_uint64), ("jobs_in_steps", C.c_uint64), ("handles_reused", C.c_uint64),
                ("cached_bytes", C.c_uint64)]


# Every exported entry point with its argument types (the declarations of
# include/hpt.h).
SIGNATURES = {
    "hpt_create": [C.POINTER(_CreateInfo), C.POINTER(C.c_void_p)],
    "hpt_set_scene": [C.c_void_p, C.POINTER(_SceneDesc)],
    "hpt_set_env_map": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float],
    "hpt_exr_info": [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "hpt_exr_decode_rgb": [C.c_char_p, C.c_size_t, C.c_void_p, C.c_uint32, C.c_uint32],
    "hpt_set_ground_color": [C.c_void_p, C.POINTER(C.c_float)],
    "hpt_set_sky_color": [C.c_void_p, C.POINTER(C.c_float)],
    "hpt_set_env_intensity": [C.c_void_p, C.c_float],
    "hpt_set_exposure_value": [C.c_void_p, C.c_float],
    "hpt_set_blue_noise": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32],
    "hpt_commit": [C.c_void_p],
    "hpt_render": [C.c_void_p, C.c_uint32, C.c_uint32],
    "hpt_reset_accumulation": [C.c_void_p],
    "hpt_set_frame_index": [C.c_void_p, C.c_uint64],
    "hpt_get_frame_index": [C.c_void_p, C.POINTER(C.c_uint64)],
    "hpt_read_image": [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t],
    "hpt_trace_closest": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "hpt_trace_any": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "hpt_trace_closest_device": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "hpt_wait_idle": [C.c_void_p],
    "hpt_destroy": [C.c_void_p],
    "hpt_last_error": [],
    "hpt_get_bvh_info": [C.c_void_p, C.POINTER(_BvhInfo)],
    "hpt_export_bvh": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t],
    "hpt_trace_counts": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p],
    "hpt_set_profiling": [C.c_void_p, C.c_int],
    "hpt_get_stage_stats": [C.c_void_p, C.c_int, C.POINTER(_StageStats)],
    "hpt_set_frames_in_flight": [C.c_void_p, C.c_uint32],
    "hpt_set_textures": [C.c_void_p, C.POINTER(_TextureDesc), C.c_uint32],
    "hpt_share_scene": [C.c_void_p, C.c_void_p],
    "hpt_query_idle": [C.c_void_p, C.POINTER(C.c_int)],
    "hpt_get_memory_usage": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "hpt_get_frames_in_flight": [C.c_void_p, C.POINTER(C.c_uint32)],
    "hpt_get_image_pointer": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)],
    "hpt_set_launch_mode": [C.c_void_p, C.c_uint32],
    "hpt_set_shade_sort": [C.c_void_p, C.c_uint32],
    "hpt_set_vmis_phases": [C.c_void_p, C.c_uint32],
    "hpt_group_create": [C.c_int32, C.POINTER(C.c_void_p)],
    "hpt_group_render": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                         C.c_uint32],
    "hpt_group_set_launch_mode": [C.c_void_p, C.c_uint32],
    "hpt_group_set_shade_sort": [C.c_void_p, C.c_uint32],
    "hpt_group_get_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "hpt_group_wait_idle": [C.c_void_p],
    "hpt_group_destroy": [C.c_void_p],
    "hpt_set_bvh_builder": [C.c_void_p, C.c_uint32],
    "hpt_get_bvh_build_stats": [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)],
    "hpt_get_launch_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "hpt_get_truncated_paths": [C.c_void_p, C.POINTER(C.c_uint64)],
    "hpt_set_tail_from": [C.c_void_p, C.c_uint32],
    "hpt_set_traversal_block": [C.c_void_p, C.c_uint32],
    "hpt_group_set_traversal_block": [C.c_void_p, C.c_uint32],
    "hpt_set_traversal_refill": [C.c_void_p, C.c_uint32],
    "hpt_group_set_traversal_refill": [C.c_void_p, C.c_uint32],
    "hpt_group_set_tail_from": [C.c_void_p, C.c_uint32],
    "hpt_group_get_truncated_paths": [C.c_void_p, C.POINTER(C.c_uint64)],
    "hpt_present_create": [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)],
    "hpt_present_frame": [C.c_void_p],
    "hpt_present_acquire": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)],
    "hpt_present_get_stats": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "hpt_present_destroy": [C.c_void_p],
    "hpt_get_image_size": [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "hpt_denoiser_create": [C.c_char_p, C.c_size_t, C.c_int32, C.POINTER(C.c_void_p)],
    "hpt_denoise": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
    "hpt_denoiser_destroy": [C.c_void_p],
    "hpt_group_set_vmis_phases": [C.c_void_p, C.c_uint32],
    "hpt_server_config_init": [C.POINTER(_ServerConfig)],
    "hpt_job_desc_init": [C.POINTER(_JobDesc)],
    "hpt_server_create": [C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)],
    "hpt_server_add_scene": [C.c_void_p, C.POINTER(_ServeSceneDesc), C.POINTER(C.c_uint32)],
    "hpt_server_set_denoiser": [C.c_void_p, C.c_void_p],
    "hpt_server_plan_job": [C.c_void_p, C.POINTER(_JobDesc), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
    "hpt_server_submit": [C.c_void_p, C.POINTER(_JobDesc), C.POINTER(C.c_uint64)],
    "hpt_server_wait": [C.c_void_p, C.c_uint64, C.POINTER(_JobResult)],
    "hpt_server_next_result": [C.c_void_p, C.c_double, C.POINTER(_JobResult), C.POINTER(C.c_int)],
    "hpt_server_read_image": [C.c_void_p, C.c_uint64, C.c_void_p, C.c_size_t],
    "hpt_server_drain": [C.c_void_p],
    "hpt_server_clock": [C.c_void_p, C.POINTER(C.c_double)],
    "hpt_server_get_stats": [C.c_void_p, C.POINTER(_ServerStats)],
    "hpt_server_destroy": [C.c_void_p],
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libhpt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        if name in ("hpt_destroy", "hpt_group_destroy", "hpt_present_destroy", "hpt_denoiser_destroy",
                    "hpt_server_destroy", "hpt_server_config_init", "hpt_job_desc_init"):
            fn.restype = None
        elif name == "hpt_last_error":
            fn.restype = C.c_char_p
        else:
            fn.restype = C.c_int32
    _lib = lib
    return lib


class HptError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _check(status: int) -> None:
    if status != 0:
        raise HptError(status, _lib.hpt_last_error().decode())


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data) if a.size else C.c_void_p(0)


# ------------------------------------------------------------------ scenes
@dataclass
class Primitive:
    """One mesh: vertices (N, 11) float32, indices (M, 3) uint32, a 4x4
    world transform (row-major numpy; passed to the ABI column-major)."""
    vertices: np.ndarray
    indices: np.ndarray
    material_index: int
    transform: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))


@dataclass
class Scene:
    cameras: np.ndarray          # CAMERA_DTYPE
    lights: np.ndarray           # LIGHT_DTYPE
    materials: np.ndarray        # MATERIAL_DTYPE
    primitives: list             # [Primitive]
    env_rgb: np.ndarray | None = None   # (H, W, 3) float32, row 0 = +Z
    env_rotation: float = 0.0
    ground_color: tuple = (0.5, 0.5, 0.5)
    sky_color: tuple = (0.5, 0.5, 0.5)
    env_intensity: float = 1.0
    textures: list = field(default_factory=list)  # RGBA8 (H, W, 4) arrays the *_map_index fields index

    @property
    def triangle_count(self) -> int:
        return int(sum(len(p.indices) for p in self.primitives))

    def to_desc(self):
        """(hpt_scene_desc, keep-alive list) for the C ABI."""
        keep = []
        cams = np.ascontiguousarray(self.cameras, CAMERA_DTYPE)
        lights = np.ascontiguousarray(self.lights, LIGHT_DTYPE)
        mats = np.ascontiguousarray(self.materials, MATERIAL_DTYPE)
        keep += [cams, lights, mats]
        prims = (_PrimitiveDesc * max(1, len(self.primitives)))()
        for i, p in enumerate(self.primitives):
            v = np.ascontiguousarray(p.vertices, np.float32).reshape(-1, VERTEX_FLOATS)
            ix = np.ascontiguousarray(p.indices, np.uint32).reshape(-1)
            t = np.ascontiguousarray(np.asarray(p.transform, np.float32).T).reshape(-1)  # column-major
            keep += [v, ix, t]
            prims[i].transform[:] = t.tolist()
            prims[i].material_index = int(p.material_index) & 0xFFFFFFFF
            prims[i].vertex_count = v.shape[0]
            prims[i].index_count = ix.shape[0]
            prims[i].vertices = v.ctypes.data
            prims[i].indices = ix.ctypes.data
        keep.append(prims)
        d = _SceneDesc(cams.ctypes.data, len(cams), len(lights), lights.ctypes.data if len(lights) else None,
                       mats.ctypes.data if len(mats) else None, len(mats), len(self.primitives),
                       C.cast(prims, C.c_void_p) if self.primitives else None)
        return d, keep

    def world_triangles(self) -> np.ndarray:
        """(T, 3, 3) float32 world-space vertices in global triangle order,
        transformed like the library does (column-major, summed left to right)."""
        out = []
        for p in self.primitives:
            v = np.asarray(p.vertices, np.float32).reshape(-1, VERTEX_FLOATS)[:, :3]
            m = np.asarray(p.transform, np.float32)
            x, y, z = v[:, 0], v[:, 1], v[:, 2]
            w = np.stack([((m[r, 0] * x + m[r, 1] * y) + m[r, 2] * z) + m[r, 3] for r in range(3)], axis=1)
            out.append(w[np.asarray(p.indices, np.int64).reshape(-1, 3)])
        return np.concatenate(out).astype(np.float32) if out else np.zeros((0, 3, 3), np.float32)


def make_rays(origins, directions, tmin=0.0, tmax=WORLD_SIZE) -> np.ndarray:
    o = np.asarray(origins, np.float32).reshape(-1, 3)
    d = np.asarray(directions, np.float32).reshape(-1, 3)
    r = np.zeros(len(o), RAY_DTYPE)
    r["origin"] = o
    r["direction"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


# ---------------------------------------------------------------- renderer
class HalaRenderer:
    """The renderer the reference host creates at hala-pathtracer.rs:128-138.

    The Vulkan-only calls (push_*_shader_with_file, the imgui context) have no
    counterpart; the shader permutation is the `features` bit set
    (hala-pathtracer.rs:101-114 chooses it from the scene).
    """

    def __init__(self, width: int, height: int, max_depth: int, rr_depth: int, enable_tonemap: bool = True,
                 enable_aces: bool = True, use_simple_aces: bool = False, max_samples: int = 0,
                 features: int = 0, exposure_value: float = 0.0, device: int = 0):
        lib = load_library()
        info = _CreateInfo(width, height, max_depth, rr_depth, int(enable_tonemap), int(enable_aces),
                           int(use_simple_aces), features, max_samples, exposure_value, device)
        h = C.c_void_p()
        _check(lib.hpt_create(C.byref(info), C.byref(h)))
        self._h = h
        self._lib = lib
        self.width, self.height = width, height

    # -- set-up (hala-pathtracer.rs:195-213)
    def set_scene(self, scene: Scene) -> None:
        """set_scene (hala-pathtracer.rs:198): the scene's arrays, its
        texture table and its environment."""
        d, keep = scene.to_desc()
        _check(self._lib.hpt_set_scene(self._h, C.byref(d)))
        del keep
        self.set_textures(scene.textures)
        if scene.env_rgb is not None:
            self.set_envmap(scene.env_rgb, scene.env_rotation)
        else:
            self.set_ground_color(scene.ground_color)
            self.set_sky_color(scene.sky_color)
        self.set_env_intensity(scene.env_intensity)

    def share_scene(self, src: "HalaRenderer") -> None:
        """Renders src's committed scene (shared device copy); commits self."""
        _check(self._lib.hpt_share_scene(self._h, src._h))

    def is_idle(self) -> bool:
        v = C.c_int()
        _check(self._lib.hpt_query_idle(self._h, C.byref(v)))
        return bool(v.value)

    def memory_usage(self) -> tuple:
        """(frame bytes, scene bytes) of device memory."""
        f, sc = C.c_uint64(), C.c_uint64()
        _check(self._lib.hpt_get_memory_usage(self._h, C.byref(f), C.byref(sc)))
        return f.value, sc.value

    def set_textures(self, textures) -> None:
        arr, keep = texture_descs(textures)
        _check(self._lib.hpt_set_textures(self._h, arr, len(keep)))

    def set_envmap(self, rgb: np.ndarray, rotation: float = 0.0) -> None:
        a = np.ascontiguousarray(rgb, np.float32)
        h, w = a.shape[:2]
        _check(self._lib.hpt_set_env_map(self._h, _ptr(a), w, h, rotation))

    def set_ground_color(self, rgb) -> None:
        c = (C.c_float * 4)(*[float(x) for x in rgb][:3], 1.0)
        _check(self._lib.hpt_set_ground_color(self._h, c))

    def set_sky_color(self, rgb) -> None:
        c = (C.c_float * 4)(*[float(x) for x in rgb][:3], 1.0)
        _check(self._lib.hpt_set_sky_color(self._h, c))

    def set_env_intensity(self, v: float) -> None:
        _check(self._lib.hpt_set_env_intensity(self._h, v))

    def set_exposure_value(self, ev: float) -> None:
        _check(self._lib.hpt_set_exposure_value(self._h, ev))

    def load_blue_noise_texture(self, rgba8: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba8, np.uint8)
        _check(self._lib.hpt_set_blue_noise(self._h, _ptr(a), a.shape[1], a.shape[0]))

    def commit(self) -> None:
        _check(self._lib.hpt_commit(self._h))

    # -- per frame (hala-pathtracer.rs:264-285)
    def render(self, frames: int = 1, camera_index: int = 0) -> None:
        _check(self._lib.hpt_render(self._h, frames, camera_index))

    def reset_accumulation(self) -> None:
        _check(self._lib.hpt_reset_accumulation(self._h))

    @property
    def frame_index(self) -> int:
        v = C.c_uint64()
        _check(self._lib.hpt_get_frame_index(self._h, C.byref(v)))
        return v.value

    @frame_index.setter
    def frame_index(self, f: int) -> None:
        _check(self._lib.hpt_set_frame_index(self._h, f))

    def render_single_frame(self, frame: int) -> np.ndarray:
        """Radiance of frame `frame` alone: clears the accumulators, renders
        that frame and undoes the running-mean weight 1/(frame+1)."""
        self.reset_accumulation()
        self.frame_index = frame
        self.render(1)
        return self.read_image(IMAGE_ACCUM)[..., :3] * np.float32(frame + 1)

    def read_image(self, which: int) -> np.ndarray:
        if which == IMAGE_FINAL_RGBA8:
            out = np.empty((self.height, self.width, 4), np.uint8)
        else:
            out = np.empty((self.height, self.width, 4), np.float32)
        _check(self._lib.hpt_read_image(self._h, which, _ptr(out), out.nbytes))
        return out

    def image_pointer(self, which: int) -> tuple:
        """(device address, bytes) of a resident image; read it only after
        wait_idle()."""
        p, n = C.c_void_p(), C.c_size_t()
        _check(self._lib.hpt_get_image_pointer(self._h, which, C.byref(p), C.byref(n)))
        return int(p.value or 0), int(n.value)

    def save_images(self, prefix: str) -> None:
        """HalaRenderer::save_images (hala-pathtracer.rs:251): colour, albedo
        and normal as PFM, the names denoise.py:11,18-19 looks for."""
        for which, suffix in ((IMAGE_ACCUM, "color"), (IMAGE_ALBEDO, "albedo"), (IMAGE_NORMAL, "normal")):
            write_pfm(f"{prefix}_{suffix}.pfm", self.read_image(which)[..., :3])

    def wait_idle(self) -> None:
        _check(self._lib.hpt_wait_idle(self._h))

    # -- ray queries and introspection
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        _check(self._lib.hpt_trace_closest(self._h, _ptr(rays), len(rays), _ptr(hits)))
        return hits

    def trace_closest_device(self, rays_ptr: int, count: int, hits_ptr: int) -> None:
        """hpt_trace_closest_device: `count` rays at device address rays_ptr
        (RAY_DTYPE) into hpt_hit records at hits_ptr, queued on the handle's
        stream with no allocation (wait_idle before reading)."""
        _check(self._lib.hpt_trace_closest_device(self._h, C.c_void_p(rays_ptr), count, C.c_void_p(hits_ptr)))

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint32)
        _check(self._lib.hpt_trace_any(self._h, _ptr(rays), len(rays), _ptr(occ)))
        return occ

    def trace_counts(self, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        nv = np.zeros(len(rays), np.uint32)
        nt = np.zeros(len(rays), np.uint32)
        _check(self._lib.hpt_trace_counts(self._h, _ptr(rays), len(rays), int(any_hit), _ptr(nv), _ptr(nt)))
        return nv, nt

    def bvh_info(self) -> dict:
        info = _BvhInfo()
        _check(self._lib.hpt_get_bvh_info(self._h, C.byref(info)))
        return {"node_count": info.node_count, "triangle_count": info.triangle_count, "max_depth": info.max_depth}

    def export_bvh(self):
        info = self.bvh_info()
        nodes = np.zeros(info["node_count"] * 80, np.uint8)
        tris = np.zeros(info["triangle_count"] * 48, np.uint8)
        _check(self._lib.hpt_export_bvh(self._h, _ptr(nodes), nodes.nbytes, _ptr(tris), tris.nbytes))
        return nodes, tris

    @property
    def frames_in_flight(self) -> int:
        """Frames traced together per hpt_render batch (the images do not
        depend on it)."""
        v = C.c_uint32()
        _check(self._lib.hpt_get_frames_in_flight(self._h, C.byref(v)))
        return v.value

    @frames_in_flight.setter
    def frames_in_flight(self, k: int) -> None:
        _check(self._lib.hpt_set_frames_in_flight(self._h, k))

    def set_bvh_builder(self, builder: int) -> None:
        """BVH_GPU (the default: binned SAH and 8-wide collapse on the device)
        or BVH_HOST; both build the same bytes.  Applies at the next commit."""
        _check(self._lib.hpt_set_bvh_builder(self._h, builder))

    def bvh_build_stats(self) -> tuple:
        """(builder used by the last commit, its wall time in ms)."""
        b, ms = C.c_uint32(), C.c_double()
        _check(self._lib.hpt_get_bvh_build_stats(self._h, C.byref(b), C.byref(ms)))
        return b.value, ms.value

    def set_launch_mode(self, mode: int) -> None:
        """LAUNCH_GRAPH (the default: full batches replayed as captured HIP
        graphs) or LAUNCH_EAGER (every kernel launched); same images."""
        _check(self._lib.hpt_set_launch_mode(self._h, mode))

    def set_vmis_phases(self, on: bool = True) -> None:
        """Phased (default) or single-kernel volume-MIS shading; same images."""
        _check(self._lib.hpt_set_vmis_phases(self._h, int(bool(on))))

    def set_shade_sort(self, first_bounce: int = SHADE_SORT_OFF) -> None:
        """Group each bounce's paths by hit material before shading, from
        bounce `first_bounce` on (SHADE_SORT_OFF: never); same images."""
        _check(self._lib.hpt_set_shade_sort(self._h, first_bounce))

    def launch_stats(self) -> tuple:
        """(batches replayed as graphs, graphs captured)."""
        a, b = C.c_uint64(), C.c_uint64()
        _check(self._lib.hpt_get_launch_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tail_from(self, first_bounce: int = TAIL_FROM_NONE) -> None:
        """Run the paths still live at bounce `first_bounce` to their end in
        one kernel, a lane per path (TAIL_FROM_NONE: every bounce up to
        max_depth + 1 in the wavefront); same images."""
        _check(self._lib.hpt_set_tail_from(self._h, first_bounce))

    def set_traversal_block(self, threads: int = 128) -> None:
        """Threads per block of the traversal kernels (extension, shadow,
        transmittance walks): 64, 128 or 256; same images."""
        _check(self._lib.hpt_set_traversal_block(self._h, threads))

    def set_traversal_refill(self, on: bool = True) -> None:
        """Lane refill in the traversal kernels (on by default): resident
        blocks whose waves take new queue entries for lanes whose rays have
        ended; same images."""
        _check(self._lib.hpt_set_traversal_refill(self._h, int(on)))

    def truncated_paths(self) -> int:
        """Paths stopped at the 4096-bounce limit since the handle was made
        (TRANSPARENT pass-throughs do not count toward max_depth)."""
        v = C.c_uint64()
        _check(self._lib.hpt_get_truncated_paths(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, flags: int) -> None:
        _check(self._lib.hpt_set_profiling(self._h, flags))

    def stage_stats(self, stage: int) -> dict:
        s = _StageStats()
        _check(self._lib.hpt_get_stage_stats(self._h, stage, C.byref(s)))
        return {"launches": s.launches, "total_ms": s.total_ms, "rays": s.rays, "node_visits": s.node_visits,
                "tri_tests": s.tri_tests, "queries": s.queries}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.hpt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_pfm(path: str, rgb: np.ndarray) -> None:
    """Little-endian colour PFM, bottom row first."""
    a = np.ascontiguousarray(np.flipud(np.asarray(rgb, np.float32)))
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (a.shape[1], a.shape[0]))
        f.write(a.astype("<f4").tobytes())



class HalaGroup:
    """Several jobs of one scene traced as one wavefront per step
    (hpt_group_*): `render(jobs, frames)` renders frames[i] frames of
    jobs[i] (HalaRenderer handles sharing one committed scene) with one
    extension, shading and shadow launch per bounce for all of them; every
    job ends as if it had rendered alone."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.hpt_group_create(device, C.byref(h)))
        self._h = h

    def render(self, jobs, frames, cameras=None) -> None:
        n = len(jobs)
        hs = (C.c_void_p * n)(*[j._h for j in jobs])
        fr = (C.c_uint32 * n)(*[int(f) for f in frames])
        cam = (C.c_uint32 * n)(*[int(c) for c in cameras]) if cameras is not None else None
        _check(self._lib.hpt_group_render(self._h, hs, fr, cam, n))

    def set_launch_mode(self, mode: int) -> None:
        _check(self._lib.hpt_group_set_launch_mode(self._h, mode))

    def set_vmis_phases(self, on: bool = True) -> None:
        """Phased (default) or single-kernel volume-MIS shading; same images."""
        _check(self._lib.hpt_group_set_vmis_phases(self._h, int(bool(on))))

    def set_shade_sort(self, first_bounce: int = SHADE_SORT_OFF) -> None:
        _check(self._lib.hpt_group_set_shade_sort(self._h, first_bounce))

    def stats(self) -> dict:
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self._lib.hpt_group_get_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"steps": a.value, "graph_launches": b.value, "device_bytes": c.value}

    def set_tail_from(self, first_bounce: int = TAIL_FROM_NONE) -> None:
        _check(self._lib.hpt_group_set_tail_from(self._h, first_bounce))

    def set_traversal_block(self, threads: int = 128) -> None:
        _check(self._lib.hpt_group_set_traversal_block(self._h, threads))

    def set_traversal_refill(self, on: bool = True) -> None:
        _check(self._lib.hpt_group_set_traversal_refill(self._h, int(on)))

    def truncated_paths(self) -> int:
        """Paths of this group's steps stopped at the 4096-bounce limit."""
        v = C.c_uint64()
        _check(self._lib.hpt_group_get_truncated_paths(self._h, C.byref(v)))
        return v.value

    def wait_idle(self) -> None:
        _check(self._lib.hpt_group_wait_idle(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.hpt_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

class Denoiser:
    """hpt_denoiser_*: the trained U-Net loaded from its .npz bytes behind the
    C ABI, denoising a renderer's resident images (the oidnDenoise step of
    denoise.py:26-30 that the reference's Save button feeds)."""

    def __init__(self, npz: bytes | str, device: int = 0):
        self._lib = load_library()
        if isinstance(npz, str):
            with open(npz, "rb") as f:
                npz = f.read()
        h = C.c_void_p()
        _check(self._lib.hpt_denoiser_create(npz, len(npz), device, C.byref(h)))
        self._h = h

    def denoise(self, r: "HalaRenderer") -> np.ndarray:
        """(H, W, 4) float32 denoised colour of the renderer's images."""
        out = np.empty((r.height, r.width, 4), np.float32)
        _check(self._lib.hpt_denoise(self._h, r._h, _ptr(out), out.nbytes, 0))
        return out

    def denoise_into(self, r: "HalaRenderer", device_ptr: int) -> None:
        """The same into a device buffer of width*height*16 bytes."""
        _check(self._lib.hpt_denoise(self._h, r._h, C.c_void_p(device_ptr), r.width * r.height * 16, 1))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.hpt_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SERVE_STREAMS = 0
SERVE_WAVEFRONT = 1
SERVE_DEFAULT = 0xFFFFFFFF


class RenderServer:
    """hpt_server_*: the serving loop behind the C ABI.  Jobs over resident
    scenes are admitted in submission order while their device footprints fit
    `budget_bytes`, run on streams of their own (SERVE_STREAMS) or in one
    wavefront per scene (SERVE_WAVEFRONT), and end with their image in host
    memory.  `dry_run_ns_per_path` > 0 schedules without a device (each batch
    sleeps frames * pixels * that many ns; images hold the job id)."""

    def __init__(self, budget_bytes: int, mode: int = SERVE_STREAMS, max_frames: int = 32, keep_images: bool = False,
                 step_paths: int = 8 << 20, launch: int = LAUNCH_GRAPH, device: int = 0, dry_run_ns_per_path: int = 0,
                 max_running: int = 8, reuse_handles: bool = True):
        self._lib = load_library()
        cfg = _ServerConfig()
        self._lib.hpt_server_config_init(C.byref(cfg))
        cfg.device, cfg.mode, cfg.budget_bytes, cfg.max_frames = device, mode, budget_bytes, max_frames
        cfg.keep_images, cfg.step_paths, cfg.launch_mode = int(keep_images), step_paths, launch
        cfg.dry_run_ns_per_path = dry_run_ns_per_path
        cfg.max_running = max_running
        cfg.reuse_handles = int(reuse_handles)
        h = C.c_void_p()
        _check(self._lib.hpt_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = []  # scenes and denoiser the server borrows

    def add_scene(self, template: "HalaRenderer | None", features: int, launch: int = SERVE_DEFAULT,
                  shade_sort: int = SERVE_DEFAULT, tail_from: int = SERVE_DEFAULT, traversal_block: int = SERVE_DEFAULT,
                  dry_run_fail: bool = False) -> int:
        d = _ServeSceneDesc(template._h if template is not None else None, features, launch, shade_sort, tail_from,
                            traversal_block, int(dry_run_fail))
        sid = C.c_uint32()
        _check(self._lib.hpt_server_add_scene(self._h, C.byref(d), C.byref(sid)))
        self._keep.append(template)
        return sid.value

    def set_denoiser(self, d: "Denoiser | None") -> None:
        _check(self._lib.hpt_server_set_denoiser(self._h, d._h if d is not None else None))
        self._keep.append(d)

    def job_desc(self, scene: int, width: int, height: int, spp: int, max_depth: int, rr_depth: int = 3,
                 **fields) -> _JobDesc:
        d = _JobDesc()
        self._lib.hpt_job_desc_init(C.byref(d))
        d.scene, d.width, d.height, d.spp, d.max_depth, d.rr_depth = scene, width, height, spp, max_depth, rr_depth
        for k, v in fields.items():
            setattr(d, k, v)
        return d

    def plan(self, desc: _JobDesc) -> tuple:
        """(footprint bytes, frames per batch or step) of a job."""
        fp, k = C.c_uint64(), C.c_uint32()
        _check(self._lib.hpt_server_plan_job(self._h, C.byref(desc), C.byref(fp), C.byref(k)))
        return fp.value, k.value

    def submit(self, desc: _JobDesc) -> int:
        jid = C.c_uint64()
        _check(self._lib.hpt_server_submit(self._h, C.byref(desc), C.byref(jid)))
        return jid.value

    @staticmethod
    def _result(r: _JobResult) -> dict:
        return {"job_id": r.job_id, "status": r.status, "width": r.width, "height": r.height, "frames": r.frames,
                "has_image": bool(r.has_image), "footprint_bytes": r.footprint_bytes, "submitted": r.submitted_s,
                "admitted": r.admitted_s, "finished": r.finished_s, "error": r.error.decode() or None}

    def wait(self, job_id: int) -> dict:
        r = _JobResult()
        _check(self._lib.hpt_server_wait(self._h, job_id, C.byref(r)))
        return self._result(r)

    def next_result(self, timeout: float = -1.0) -> dict | None:
        r, got = _JobResult(), C.c_int()
        _check(self._lib.hpt_server_next_result(self._h, timeout, C.byref(r), C.byref(got)))
        return self._result(r) if got.value else None

    def read_image(self, job_id: int, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.float32)
        _check(self._lib.hpt_server_read_image(self._h, job_id, _ptr(out), out.nbytes))
        return out

    def drain(self) -> None:
        _check(self._lib.hpt_server_drain(self._h))

    def clock(self) -> float:
        t = C.c_double()
        _check(self._lib.hpt_server_clock(self._h, C.byref(t)))
        return t.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(self._lib.hpt_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.hpt_server_destroy(self._h)
            self._h = None
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Presenter:
    """hpt_present_*: a ring of pinned host copies of a renderer's final
    rgba8 image for a window loop (the reference presents through its
    swapchain, hala-pathtracer.rs:236-289).  present() after render() queues
    a copy without waiting; acquire() returns the newest finished copy and
    its frame count, or None."""

    def __init__(self, renderer: "HalaRenderer", slots: int = 3):
        self._lib = load_library()
        self.renderer = renderer
        h = C.c_void_p()
        _check(self._lib.hpt_present_create(renderer._h, slots, C.byref(h)))
        self._h = h

    def present(self) -> None:
        _check(self._lib.hpt_present_frame(self._h))

    def acquire(self):
        """(frame count, (H, W, 4) uint8 copy of the slot) or None."""
        p, f = C.c_void_p(), C.c_uint64()
        _check(self._lib.hpt_present_acquire(self._h, C.byref(p), C.byref(f)))
        if not p.value:
            return None
        r = self.renderer
        buf = (C.c_uint8 * (r.width * r.height * 4)).from_address(p.value)
        return f.value, np.frombuffer(buf, np.uint8).reshape(r.height, r.width, 4).copy()

    def stats(self) -> tuple:
        a, b = C.c_uint64(), C.c_uint64()
        _check(self._lib.hpt_present_get_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.hpt_present_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_exr(data: bytes) -> np.ndarray:
    """An OpenEXR file's bytes as (H, W, 3) float32 RGB, row 0 at the top
    (hpt_exr_info + hpt_exr_decode_rgb; host code, no device)."""
    lib = load_library()
    w, h = C.c_uint32(), C.c_uint32()
    _check(lib.hpt_exr_info(data, len(data), C.byref(w), C.byref(h)))
    out = np.empty((h.value, w.value, 3), np.float32)
    _check(lib.hpt_exr_decode_rgb(data, len(data), _ptr(out), w.value, h.value))
    return out


def read_exr(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        return decode_exr(f.read())


def read_pfm(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return np.flipud(data).astype(np.float32)
