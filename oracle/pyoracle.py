"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "hala-pathtracer_amd"))
import hpt  # noqa: E402  (scene layouts and dtypes)

_lib = None


class Settings(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("enable_tonemap", C.c_uint32), ("enable_aces", C.c_uint32), ("use_simple_aces", C.c_uint32),
                ("features", C.c_uint32), ("camera_index", C.c_uint32), ("exposure_value", C.c_float)]


class Frame(C.Structure):
    _fields_ = [("tangent", C.c_float * 3), ("bitangent", C.c_float * 3), ("n", C.c_float * 3),
                ("eta", C.c_float), ("depth", C.c_uint32), ("any_non_specular_bounce", C.c_uint32)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sig = {
            "or_scene_create": ([vp], vp), "or_scene_destroy": ([vp], None),
            "or_scene_set_env_map": ([vp, vp, u32, u32, f32], i32),
            "or_scene_set_env_sky": ([vp, vp, vp], None), "or_scene_set_env_intensity": ([vp, f32], None),
            "or_scene_set_blue_noise": ([vp, vp, u32, u32], i32), "or_scene_triangle_count": ([vp], u32),
            "or_scene_set_textures": ([vp, vp, u32], i32),
            "or_texture_sample": ([vp, u32, f32, f32, i32, vp], f32),
            "or_trace_closest_brute": ([vp, vp, u32, vp], None), "or_trace_closest": ([vp, vp, u32, vp], None),
            "or_trace_any": ([vp, vp, u32, vp], None), "or_trace_any_brute": ([vp, vp, u32, vp], None),
            "or_trace_closest_fp64": ([vp, vp, u32, vp, vp, vp], None),
            "or_trace_bvh8": ([vp, vp, u32, vp, u32, vp, u32, i32, vp, vp, vp], None),
            "or_camera_ray": ([vp, vp, u32, u32, u32, vp], None),
            "or_trace_paths": ([vp, vp, u32, u32, u32, u32, u32, vp, vp, vp, i32], i32),
            "or_render": ([vp, vp, u32, u32, vp, vp, vp, vp, i32], i32),
            "or_final_color": ([vp, vp, vp], None),
            "or_pcg4d": ([vp], None), "or_u32_to_unit": ([u32], f32),
            "or_rng_init": ([u32, u32, u32, vp, vp], None), "or_blue_noise_texel": ([vp, u32, u32, vp], None),
            "or_bxdf_eval": ([vp, vp, vp, vp, vp, vp], None),
            "or_bxdf_sample": ([vp, vp, vp, vp, vp, vp, vp, vp], None),
            "or_light_sample": ([vp, u32, vp, vp, vp, vp, vp, vp, vp], None),
            "or_light_intersect": ([vp, vp, vp, vp], f32),
            "or_build_alias": ([vp, u32, vp, vp], C.c_double),
            "or_env_sample": ([vp, vp, vp, vp, vp], None), "or_env_eval": ([vp, vp, u32, f32, vp], None),
            "or_env_pdf_dir": ([vp, vp], f32), "or_env_alias": ([vp, vp, vp, vp, vp], None),
            "or_sample_hg": ([vp, f32, f32, f32, vp], None), "or_phase_hg": ([f32, f32], f32),
            "or_sample_ggx_vndf": ([vp, f32, f32, f32, f32, vp], None),
            "or_cosine_sample_hemisphere": ([f32, f32, vp], None),
            "or_sample_gtr1": ([f32, f32, f32, vp], None),
            "or_tonemap_aces_fitted": ([vp, vp], None), "or_tonemap_aces": ([vp, vp], None),
            "or_tonemap_reinhard": ([vp, f32, vp], None), "or_linear_2_srgb": ([f32], f32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def P(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a.size else None


def f3(x):
    """A ctypes float[3]: passed directly as an argument it stays alive for
    the duration of the call (a numpy temporary behind a raw address would
    not)."""
    return (C.c_float * 3)(*[float(v) for v in np.asarray(x, np.float64).reshape(3)])


class OracleScene:
    """The oracle's copy of a scene (same hpt_scene_desc as the product)."""

    def __init__(self, scene: "hpt.Scene"):
        L = lib()
        d, keep = scene.to_desc()
        self._p = L.or_scene_create(C.byref(d))
        del keep
        if not self._p:
            raise RuntimeError("or_scene_create failed")
        if scene.env_rgb is not None:
            a = np.ascontiguousarray(scene.env_rgb, np.float32)
            if L.or_scene_set_env_map(self._p, P(a), a.shape[1], a.shape[0], scene.env_rotation) != 0:
                raise RuntimeError("or_scene_set_env_map failed")
        else:
            L.or_scene_set_env_sky(self._p, f3((scene.ground_color)), f3((scene.sky_color)))
        L.or_scene_set_env_intensity(self._p, scene.env_intensity)
        arr, keep = hpt.texture_descs(scene.textures)
        if L.or_scene_set_textures(self._p, C.cast(arr, C.c_void_p), len(keep)) != 0:
            raise RuntimeError("or_scene_set_textures failed")

    def texture_sample(self, t: int, u: float, v: float, srgb: bool) -> np.ndarray:
        out = np.zeros(4, np.float32)
        lib().or_texture_sample(self._p, t, u, v, int(srgb), P(out))
        return out

    def close(self):
        if self._p:
            lib().or_scene_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_blue_noise(self, rgba8: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba8, np.uint8)
        assert lib().or_scene_set_blue_noise(self._p, P(a), a.shape[1], a.shape[0]) == 0

    # -- ray queries
    def trace_closest(self, rays, brute=False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, hpt.RAY_DTYPE)
        hits = np.zeros(len(rays), hpt.HIT_DTYPE)
        fn = lib().or_trace_closest_brute if brute else lib().or_trace_closest
        fn(self._p, P(rays), len(rays), P(hits))
        return hits

    def trace_any(self, rays, brute=False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, hpt.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint32)
        fn = lib().or_trace_any_brute if brute else lib().or_trace_any
        fn(self._p, P(rays), len(rays), P(occ))
        return occ

    def trace_fp64(self, rays):
        rays = np.ascontiguousarray(rays, hpt.RAY_DTYPE)
        t = np.zeros(len(rays), np.float64)
        tri = np.zeros(len(rays), np.uint32)
        gap = np.zeros(len(rays), np.float64)
        lib().or_trace_closest_fp64(self._p, P(rays), len(rays), P(t), P(tri), P(gap))
        return t, tri, gap

    def trace_bvh8(self, nodes: np.ndarray, tris: np.ndarray, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, hpt.RAY_DTYPE)
        hits = np.zeros(len(rays), hpt.HIT_DTYPE)
        nv = np.zeros(len(rays), np.uint32)
        nt = np.zeros(len(rays), np.uint32)
        lib().or_trace_bvh8(self._p, P(nodes), len(nodes) // 80, P(tris), len(tris) // 48, P(rays), len(rays),
                            int(any_hit), P(hits), P(nv), P(nt))
        return hits, nv, nt

    # -- camera and paths
    def camera_rays(self, st: Settings, frame: int, xs, ys) -> np.ndarray:
        out = np.zeros(len(xs), hpt.RAY_DTYPE)
        one = np.zeros(1, hpt.RAY_DTYPE)
        for i, (x, y) in enumerate(zip(xs, ys)):
            lib().or_camera_ray(self._p, C.byref(st), frame, int(x), int(y), P(one))
            out[i] = one[0]
        return out

    def trace_paths(self, st: Settings, frame: int, rect=None, threads=0):
        x0, y0, x1, y1 = rect if rect else (0, 0, st.width, st.height)
        n = (x1 - x0) * (y1 - y0)
        rad = np.zeros((n, 3), np.float32)
        alb = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        assert lib().or_trace_paths(self._p, C.byref(st), frame, x0, y0, x1, y1, P(rad), P(alb), P(nrm),
                                    threads) == 0
        shape = (y1 - y0, x1 - x0, 3)
        return rad.reshape(shape), alb.reshape(shape), nrm.reshape(shape)

    def render(self, st: Settings, frames: int, frame_begin=0, images=None, threads=0):
        n = st.width * st.height
        if images is None:
            images = {k: np.zeros((st.height, st.width, 4), np.float32) for k in ("accum", "albedo", "normal")}
            images["final"] = np.zeros((st.height, st.width, 4), np.uint8)
        assert lib().or_render(self._p, C.byref(st), frame_begin, frames, P(images["accum"]), P(images["albedo"]),
                               P(images["normal"]), P(images["final"]), threads) == 0
        return images

    # -- environment
    def env_alias(self, n):
        prob = np.zeros(n, np.float32)
        alias = np.zeros(n, np.uint32)
        lum = np.zeros(n, np.float32)
        total = C.c_float()
        lib().or_env_alias(self._p, P(prob), P(alias), P(lum), C.byref(total))
        return prob, alias, lum, total.value

    def env_sample(self, rng4):
        rng = np.ascontiguousarray(np.asarray(rng4, np.uint32).reshape(4))
        d = np.zeros(3, np.float32)
        c = np.zeros(3, np.float32)
        pdf = C.c_float()
        lib().or_env_sample(self._p, P(rng), P(d), P(c), C.byref(pdf))
        return d, c, pdf.value, rng

    def env_eval(self, direction, depth=0, bsdf_pdf=1.0):
        out = np.zeros(3, np.float32)
        lib().or_env_eval(self._p, f3((direction)), depth, bsdf_pdf, P(out))
        return out

    def env_pdf(self, direction) -> float:
        return lib().or_env_pdf_dir(self._p, f3((direction)))

    def blue_noise_texel(self, tx, ty):
        out = np.zeros(4, np.uint8)
        lib().or_blue_noise_texel(self._p, tx, ty, P(out))
        return out


def settings(width, height, max_depth, rr_depth, features=0, camera_index=0, tonemap=True, aces=True,
             simple_aces=False, exposure=0.0) -> Settings:
    return Settings(width, height, max_depth, rr_depth, int(tonemap), int(aces), int(simple_aces), features,
                    camera_index, exposure)


# ---------------------------------------------------------------- functions
def pcg4d(v):
    a = np.ascontiguousarray(np.asarray(v, np.uint32).reshape(4).copy())
    lib().or_pcg4d(P(a))
    return a


def u32_to_unit(u: int) -> float:
    return lib().or_u32_to_unit(u)


def rng_init(x, y, frame):
    s0 = np.zeros(4, np.uint32)
    s1 = np.zeros(4, np.uint32)
    lib().or_rng_init(x, y, frame, P(s0), P(s1))
    return s0, s1


def make_frame(n, tangent, bitangent, eta=1.0, depth=0, anb=False) -> Frame:
    fr = Frame()
    fr.tangent[:] = list(map(float, tangent))
    fr.bitangent[:] = list(map(float, bitangent))
    fr.n[:] = list(map(float, n))
    fr.eta = eta
    fr.depth = depth
    fr.any_non_specular_bounce = int(anb)
    return fr


def bxdf_eval(mat, fr: Frame, V, L):
    m = np.ascontiguousarray(np.asarray(mat, hpt.MATERIAL_DTYPE).reshape(1))
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    lib().or_bxdf_eval(P(m), C.byref(fr), f3((V)), f3((L)), P(f), C.byref(pdf))
    return f, pdf.value


def bxdf_sample(mat, fr: Frame, V, rng4):
    m = np.ascontiguousarray(np.asarray(mat, hpt.MATERIAL_DTYPE).reshape(1))
    rng = np.ascontiguousarray(np.asarray(rng4, np.uint32).reshape(4).copy())
    L = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    flags = C.c_uint32()
    lib().or_bxdf_sample(P(m), C.byref(fr), f3((V)), P(rng), P(L), P(f), C.byref(pdf), C.byref(flags))
    return L, f, pdf.value, flags.value, rng


def light_sample(light, num_lights, p, rng4):
    l = np.ascontiguousarray(np.asarray(light, hpt.LIGHT_DTYPE).reshape(1))
    rng = np.ascontiguousarray(np.asarray(rng4, np.uint32).reshape(4).copy())
    d, n, e = (np.zeros(3, np.float32) for _ in range(3))
    dist, pdf = C.c_float(), C.c_float()
    lib().or_light_sample(P(l), num_lights, f3((p)), P(rng), P(d), P(n), P(e), C.byref(dist), C.byref(pdf))
    return d, n, e, dist.value, pdf.value, rng


def light_intersect(light, org, d):
    l = np.ascontiguousarray(np.asarray(light, hpt.LIGHT_DTYPE).reshape(1))
    pdf = C.c_float()
    t = lib().or_light_intersect(P(l), f3((org)), f3((d)), C.byref(pdf))
    return t, pdf.value


def build_alias(lum):
    lum = np.ascontiguousarray(lum, np.float32)
    prob = np.zeros(len(lum), np.float32)
    alias = np.zeros(len(lum), np.uint32)
    total = lib().or_build_alias(P(lum), len(lum), P(prob), P(alias))
    return prob, alias, total


def sample_hg(V, g, r1, r2):
    out = np.zeros(3, np.float32)
    lib().or_sample_hg(f3((V)), g, r1, r2, P(out))
    return out


def phase_hg(c, g):
    return lib().or_phase_hg(c, g)


def sample_ggx_vndf(wi, ax, ay, r1, r2):
    out = np.zeros(3, np.float32)
    lib().or_sample_ggx_vndf(f3((wi)), ax, ay, r1, r2, P(out))
    return out


def cosine_sample_hemisphere(r1, r2):
    out = np.zeros(3, np.float32)
    lib().or_cosine_sample_hemisphere(r1, r2, P(out))
    return out


def sample_gtr1(roughness, r1, r2):
    out = np.zeros(3, np.float32)
    lib().or_sample_gtr1(roughness, r1, r2, P(out))
    return out


def tonemap(kind, rgb, limit=1.5):
    out = np.zeros(3, np.float32)
    if kind == "aces_fitted":
        lib().or_tonemap_aces_fitted(f3((rgb)), P(out))
    elif kind == "aces":
        lib().or_tonemap_aces(f3((rgb)), P(out))
    else:
        lib().or_tonemap_reinhard(f3((rgb)), limit, P(out))
    return out


def linear_2_srgb(x):
    return lib().or_linear_2_srgb(x)


def final_color(st: Settings, rgb):
    out = np.zeros(3, np.float32)
    lib().or_final_color(C.byref(st), f3((rgb)), P(out))
    return out
