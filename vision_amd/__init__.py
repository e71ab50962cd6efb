"""vision_amd — MI355X-native megakernel path tracer behind Vision's scene format.

Python mirror of the C ABI in include/vision_amd.h (ctypes over
libvision_amd.so).  The flow follows the reference's pipeline
(src/base/mgr/pipeline.cpp): load a scene JSON, prepare (BVH build + upload),
render progressive frames, read the film.

The HIP library is required: importing works without it, but every call
raises ``LibraryMissing`` when libvision_amd.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import Optional

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libvision_amd.so"

VS_OK = 0
STATUS_NAMES = {
    0: "ok", 1: "invalid argument", 2: "parse error", 3: "unsupported", 4: "device error",
    5: "out of memory", 6: "invalid state",
}


class LibraryMissing(RuntimeError):
    pass


class VisionError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("t_min", C.c_float), ("dir", C.c_float * 3), ("t_max", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("inst", C.c_uint32), ("prim", C.c_uint32), ("bary", C.c_float * 2), ("t", C.c_float),
                ("tri", C.c_uint32)]


RAY_DTYPE = np.dtype([("org", "<f4", 3), ("t_min", "<f4"), ("dir", "<f4", 3), ("t_max", "<f4")])
HIT_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("bary", "<f4", 2), ("t", "<f4"), ("tri", "<u4")])


class SceneInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "num_triangles", "num_vertices", "num_instances", "num_materials", "num_lights", "width", "height",
        "output_spp", "max_depth", "spectrum_dim", "bvh_nodes", "bvh_depth", "frame_index", "accel_built")] + [
        ("bvh_build_ms", C.c_double), ("bvh_sah_cost", C.c_double)]


class RenderStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("rays", C.c_uint64), ("node_fetches", C.c_uint64),
                ("tri_fetches", C.c_uint64), ("launches", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("stage_ms", C.c_double * 3), ("stage_launches", C.c_uint32 * 3), ("stage_rays", C.c_uint64 * 3),
                ("stage_node_fetches", C.c_uint64 * 3), ("stage_tri_fetches", C.c_uint64 * 3)]


class SceneArrays(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("uvs", C.POINTER(C.c_float)), ("triangles", C.POINTER(C.c_uint32)),
                ("tri_instance", C.POINTER(C.c_uint32)), ("instances", C.POINTER(C.c_uint32)),
                ("materials", C.POINTER(C.c_uint32)), ("lights", C.POINTER(C.c_uint32)),
                ("camera", C.POINTER(C.c_float)), ("render", C.POINTER(C.c_float)),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_instances", C.c_uint32),
                ("num_materials", C.c_uint32), ("num_lights", C.c_uint32),
                ("spd", C.POINTER(C.c_float)), ("num_spd", C.c_uint32), ("spectrum_rgb", C.c_uint32),
                ("spectrum_dim", C.c_uint32), ("env_light", C.c_uint32), ("env_w", C.c_uint32),
                ("env_h", C.c_uint32), ("env_rgb", C.POINTER(C.c_float)), ("env_w2o", C.c_float * 9),
                ("env_o2w", C.c_float * 9), ("env_scale", C.c_float), ("world_radius", C.c_float),
                ("restir", C.POINTER(C.c_float)), ("num_restir", C.c_uint32),
                ("textures", C.POINTER(C.c_float)), ("num_texels", C.c_uint32),
                ("texture_info", C.POINTER(C.c_uint32)), ("num_textures", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_uint32), ("step_ms", C.c_float),
                ("tile_w", C.c_uint32), ("waves_per_cu", C.c_uint32), ("use_graphs", C.c_int32),
                ("count_rays", C.c_int32), ("min_waves_per_eu", C.c_uint32), ("lane_min_samples", C.c_uint64)]


class JobDesc(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("scene", "width", "height", "spp", "max_depth", "spectral",
                                           "spectrum_dim", "reserved")]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("frames_done", C.c_uint32), ("route", C.c_uint32), ("pad", C.c_uint32),
                ("samples", C.c_uint64), ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = ([(n, C.c_uint64) for n in ("jobs_submitted", "jobs_done", "jobs_failed", "samples_done", "rays",
                                           "node_fetches", "tri_fetches", "steps", "graph_launches")] +
                [(n, C.c_double) for n in ("busy_ms", "wall_ms", "msamples_per_s", "latency_p50_ms",
                                           "latency_p99_ms", "latency_mean_ms", "latency_max_ms")] +
                [(n, C.c_uint32) for n in ("max_active", "active", "queued", "pad")] +
                [("hbm_budget", C.c_uint64), ("hbm_peak", C.c_uint64), ("lane_samples", C.c_uint64),
                 ("lane_stage_ms", C.c_double * 3), ("lane_stage_launches", C.c_uint64 * 3),
                 ("lane_rays", C.c_uint64 * 2), ("lane_node_fetches", C.c_uint64 * 2),
                 ("lane_tri_fetches", C.c_uint64 * 2)])


class AutotuneResult(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("waves_per_cu", C.c_uint32), ("min_waves_per_eu", C.c_uint32),
                ("msamples_per_s", C.c_double), ("trials", C.c_uint32), ("trial_tile_w", C.c_uint32 * 16),
                ("trial_waves_per_cu", C.c_uint32 * 16), ("trial_min_waves_per_eu", C.c_uint32 * 16),
                ("trial_msamples_per_s", C.c_double * 16), ("routes", C.c_uint32),
                ("route_scene", C.c_uint32 * 16), ("route_spectral", C.c_uint32 * 16),
                ("route_choice", C.c_uint32 * 16), ("route_megakernel_ms", C.c_double * 16),
                ("route_wavefront_ms", C.c_double * 16)]


JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class NrcConfig(C.Structure):
    _fields_ = [("term_bounce", C.c_uint32), ("train_paths", C.c_uint32), ("train_steps", C.c_uint32),
                ("train_batch", C.c_uint32), ("learning_rate", C.c_float), ("min_res", C.c_uint32),
                ("max_res", C.c_uint32), ("seed", C.c_uint32)]


class NrcStats(C.Structure):
    _fields_ = ([(n, C.c_uint32) for n in ("frames", "adam_steps", "queries_last_frame", "train_samples_last_frame",
                                           "train_paths_per_frame")] +
                [(n, C.c_float) for n in ("last_loss", "render_ms", "train_paths_ms", "infer_ms", "train_ms",
                                          "accumulate_ms")])


NRC_MLP_PARAMS = 10432
NRC_GRID_PARAMS = 16 * (1 << 19) * 2
NRC_TERMINATE, NRC_TRAIN = 1, 2


# every entry point of include/vision_amd.h: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "vs_version": (C.c_int32, []),
    "vs_rgb_to_spectrum": (C.c_int32, [_P, C.c_uint32, C.c_int32, _P]),
    "vs_create": (C.c_int32, [C.c_int32, C.POINTER(_P)]),
    "vs_destroy": (None, [_P]),
    "vs_last_error": (C.c_char_p, [_P]),
    "vs_scene_load": (C.c_int32, [_P, C.c_char_p, C.c_char_p, C.POINTER(_P)]),
    "vs_scene_load_file": (C.c_int32, [_P, C.c_char_p, C.POINTER(_P)]),
    "vs_scene_destroy": (None, [_P]),
    "vs_scene_build_accel": (C.c_int32, [_P]),
    "vs_scene_info_get": (C.c_int32, [_P, C.POINTER(SceneInfo)]),
    "vs_scene_export": (C.c_int32, [_P, C.POINTER(SceneArrays)]),
    "vs_scene_bvh_export": (C.c_int32, [_P, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32),
                                        C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32)]),
    "vs_scene_set_resolution": (C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    "vs_scene_set_integrator": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_float]),
    "vs_scene_set_spectrum": (C.c_int32, [_P, C.c_int32, C.c_uint32]),
    "vs_scene_update_camera": (C.c_int32, [_P, _P, C.c_float]),
    "vs_scene_restir_state": (C.c_int32, [_P, _P, _P]),
    "vs_scene_restir_gi_state": (C.c_int32, [_P, _P, _P]),
    "vs_scene_set_denoiser": (C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "vs_scene_set_adaptive": (C.c_int32, [_P, C.c_int32, C.c_float, C.c_uint32]),
    "vs_scene_variance_stats": (C.c_int32, [_P, _P]),
    "vs_scene_svgf_state": (C.c_int32, [_P, _P, _P, _P, _P, _P]),
    "vs_trace_closest": (C.c_int32, [_P, _P, C.c_uint32, _P]),
    "vs_trace_any": (C.c_int32, [_P, _P, C.c_uint32, _P]),
    "vs_trace_closest_device": (C.c_int32, [_P, _P, C.c_uint32, _P, _P]),
    "vs_trace_any_device": (C.c_int32, [_P, _P, C.c_uint32, _P, _P]),
    "vs_trace_closest_noalloc": (C.c_int32, [_P, _P, C.c_uint32, _P]),
    "vs_scene_bsdf": (C.c_int32, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, C.c_int32, C.c_uint32, _P, _P, _P]),
    "vs_scene_set_kernel": (C.c_int32, [_P, C.c_int32, C.c_uint32, C.c_uint32]),
    "vs_render": (C.c_int32, [_P, C.c_uint32, C.c_int32]),
    "vs_set_counting": (C.c_int32, [_P, C.c_int32]),
    "vs_last_render_stats": (C.c_int32, [_P, C.POINTER(RenderStats)]),
    "vs_reset_film": (C.c_int32, [_P]),
    "vs_film_download": (C.c_int32, [_P, C.c_int32, _P]),
    "vs_film_final_picture": (C.c_int32, [_P, C.c_int32, _P]),
    "vs_film_save": (C.c_int32, [_P, C.c_char_p]),
    "vs_image_write": (C.c_int32, [_P, C.c_char_p, C.c_uint32, C.c_uint32, _P]),
    "vs_image_read": (C.c_int32, [_P, C.c_char_p, _P, _P, _P]),
    "vs_server_create": (C.c_int32, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "vs_server_add_scene": (C.c_int32, [_P, _P, C.POINTER(C.c_uint32)]),
    "vs_server_submit": (C.c_int32, [_P, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "vs_server_wait": (C.c_int32, [_P, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "vs_server_job_info": (C.c_int32, [_P, C.c_uint64, C.POINTER(JobInfo)]),
    "vs_server_result": (C.c_int32, [_P, C.c_uint64, C.c_int32, _P]),
    "vs_server_release": (C.c_int32, [_P, C.c_uint64]),
    "vs_server_drain": (C.c_int32, [_P, C.c_double]),
    "vs_server_stats_get": (C.c_int32, [_P, C.POINTER(ServerStats)]),
    "vs_server_reset_stats": (C.c_int32, [_P]),
    "vs_server_set_config": (C.c_int32, [_P, C.POINTER(ServerConfig)]),
    "vs_server_get_config": (C.c_int32, [_P, C.POINTER(ServerConfig)]),
    "vs_server_autotune": (C.c_int32, [_P, _P, C.c_uint32, C.POINTER(AutotuneResult)]),
    "vs_server_set_route": (C.c_int32, [_P, C.c_uint32, C.c_int32, C.c_uint32]),
    "vs_server_get_route": (C.c_int32, [_P, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32)]),
    "vs_server_last_error": (C.c_char_p, []),
    "vs_nrc_default_config": (None, [C.POINTER(NrcConfig)]),
    "vs_nrc_create": (C.c_int32, [_P, C.POINTER(NrcConfig), C.POINTER(_P)]),
    "vs_nrc_render": (C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "vs_nrc_stats_get": (C.c_int32, [_P, C.POINTER(NrcStats)]),
    "vs_nrc_levels": (C.c_int32, [_P, _P, _P, _P, _P]),
    "vs_nrc_get_params": (C.c_int32, [_P, _P, _P]),
    "vs_nrc_set_params": (C.c_int32, [_P, _P, _P]),
    "vs_nrc_forward": (C.c_int32, [_P, _P, C.c_uint32, _P]),
    "vs_nrc_gradients": (C.c_int32, [_P, _P, C.c_uint32, _P, _P, _P]),
    "vs_nrc_train": (C.c_int32, [_P, _P, C.c_uint32, C.c_uint32, _P]),
    "vs_nrc_last_error": (C.c_char_p, []),
    "vs_nrc_destroy": (None, [_P]),
    "vs_server_destroy": (None, [_P]),
}

_lib_handle: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libvision_amd.so (built in-tree by vision_amd/build.py)."""
    global _lib_handle
    if _lib_handle is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m vision_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def rgb_to_spectrum(rgb: np.ndarray, unbound: bool = False) -> np.ndarray:
    """Sigmoid-polynomial coefficients (c0, c1, c2, scale-or-luminance) of
    hero-wavelength mode for float32 RGB[n, 3] (hero.cpp:91-139)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((rgb.shape[0], 4), np.float32)
    st = lib().vs_rgb_to_spectrum(rgb.ctypes.data, rgb.shape[0], 1 if unbound else 0, out.ctypes.data)
    if st != 0:
        raise VisionError(st, "rgb_to_spectrum failed (spectrum table missing?)")
    return out


class Context:
    """A HIP device context (device=-1: host-only, for parsing and export)."""

    def __init__(self, device: int = 0):
        self._h = _P()
        st = lib().vs_create(device, C.byref(self._h))
        if st != VS_OK:
            raise VisionError(st, f"vs_create(device={device}) failed")
        self.device = device

    def _check(self, st: int):
        if st != VS_OK:
            raise VisionError(st, lib().vs_last_error(self._h).decode())

    def load(self, path: str | os.PathLike) -> "Scene":
        s = _P()
        self._check(lib().vs_scene_load_file(self._h, str(path).encode(), C.byref(s)))
        return Scene(self, s)

    def load_text(self, text: str | dict, base_dir: str = ".") -> "Scene":
        if isinstance(text, dict):
            text = json.dumps(text)
        s = _P()
        self._check(lib().vs_scene_load(self._h, text.encode(), str(base_dir).encode(), C.byref(s)))
        return Scene(self, s)

    def write_image(self, path: str | os.PathLike, rgba: np.ndarray):
        """Write float RGBA rows (top first) as .exr / .hdr / .png / .pfm."""
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("rgba must be float[H, W, 4]")
        self._check(lib().vs_image_write(self._h, str(path).encode(), a.shape[1], a.shape[0], a.ctypes.data))

    def read_image(self, path: str | os.PathLike) -> np.ndarray:
        """float32[H, W, 4] from .exr / .hdr / .png / .pfm (8/16-bit values as code / max)."""
        w, h = C.c_uint32(), C.c_uint32()
        self._check(lib().vs_image_read(self._h, str(path).encode(), C.byref(w), C.byref(h), None))
        out = np.zeros((h.value, w.value, 4), dtype=np.float32)
        self._check(lib().vs_image_read(self._h, str(path).encode(), C.byref(w), C.byref(h), out.ctypes.data))
        return out

    def close(self):
        if self._h:
            lib().vs_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Scene:
    """A loaded scene (the reference's Scene + Pipeline for one render)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle

    def close(self):
        if self._h:
            lib().vs_scene_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- preparation
    def build_accel(self) -> "Scene":
        self.ctx._check(lib().vs_scene_build_accel(self._h))
        return self

    def info(self) -> dict:
        i = SceneInfo()
        self.ctx._check(lib().vs_scene_info_get(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in SceneInfo._fields_}

    def set_resolution(self, w: int, h: int):
        self.ctx._check(lib().vs_scene_set_resolution(self._h, w, h))

    def set_integrator(self, max_depth: int, min_depth: int, rr_threshold: float):
        self.ctx._check(lib().vs_scene_set_integrator(self._h, max_depth, min_depth, rr_threshold))

    def set_spectrum(self, spectral: bool, dimension: int = 4):
        self.ctx._check(lib().vs_scene_set_spectrum(self._h, 1 if spectral else 0, dimension))

    def export(self) -> dict:
        """Flattened scene arrays (copies) — the input of the CPU oracle."""
        a = SceneArrays()
        self.ctx._check(lib().vs_scene_export(self._h, C.byref(a)))

        def arr(ptr, n, dtype):
            if n == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

        return {
            "positions": arr(a.positions, 3 * a.num_vertices, np.float32).reshape(-1, 3),
            "normals": arr(a.normals, 3 * a.num_vertices, np.float32).reshape(-1, 3),
            "uvs": arr(a.uvs, 2 * a.num_vertices, np.float32).reshape(-1, 2),
            "triangles": arr(a.triangles, 3 * a.num_triangles, np.uint32).reshape(-1, 3),
            "tri_instance": arr(a.tri_instance, a.num_triangles, np.uint32),
            "instances": arr(a.instances, 4 * a.num_instances, np.uint32).reshape(-1, 4),
            "materials": arr(a.materials, 32 * a.num_materials, np.uint32).reshape(-1, 32),
            "lights": arr(a.lights, 28 * a.num_lights, np.uint32).reshape(-1, 28),
            "camera": arr(a.camera, 32, np.float32),
            "render": arr(a.render, 12, np.float32),
            "spd": arr(a.spd, a.num_spd, np.float32),
            "spectrum_rgb": int(a.spectrum_rgb),
            "spectrum_dim": int(a.spectrum_dim),
            "env_light": int(a.env_light),
            "env_rgb": arr(a.env_rgb, 3 * a.env_w * a.env_h, np.float32).reshape(a.env_h, a.env_w, 3),
            "env_w2o": np.array(a.env_w2o[:], np.float32).reshape(3, 3),
            "env_o2w": np.array(a.env_o2w[:], np.float32).reshape(3, 3),
            "env_scale": float(a.env_scale),
            "world_radius": float(a.world_radius),
            "restir": arr(a.restir, a.num_restir, np.float32),
            "textures": arr(a.textures, 4 * a.num_texels, np.float32).reshape(-1, 4),
            "texture_info": arr(a.texture_info, 4 * a.num_textures, np.uint32).reshape(-1, 4),
        }

    def bvh(self) -> tuple[np.ndarray, np.ndarray]:
        """Host copy of the BVH8: (nodes uint32[N,20], leaf triangles float32[M,12])."""
        pn = C.POINTER(C.c_uint32)()
        pt = C.POINTER(C.c_float)()
        nn = C.c_uint32()
        nt = C.c_uint32()
        self.ctx._check(lib().vs_scene_bvh_export(self._h, C.byref(pn), C.byref(nn), C.byref(pt), C.byref(nt)))
        nodes = np.ctypeslib.as_array(pn, shape=(nn.value * 20,)).copy().reshape(-1, 20)
        tris = np.ctypeslib.as_array(pt, shape=(nt.value * 12,)).copy().reshape(-1, 12)
        return nodes, tris

    # -- ray queries (Geometry::trace_closest / trace_occlusion)
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        self.ctx._check(lib().vs_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def trace_closest_noalloc(self, rays: np.ndarray) -> np.ndarray:
        """trace_closest through vs_trace_closest_noalloc (no device
        allocation; at most 262144 rays per call)."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        self.ctx._check(lib().vs_trace_closest_noalloc(self._h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(len(rays), dtype=np.uint8)
        self.ctx._check(lib().vs_trace_any(self._h, rays.ctypes.data, len(rays), occ.ctypes.data))
        return occ

    # -- rendering (PathTracingIntegrator::render, one frame = one spp)
    def bsdf(self, mat: int, wo: np.ndarray, wi: np.ndarray | None = None, lam: np.ndarray | None = None,
             frame: int = 0):
        """Evaluate (wi given) or sample (wi None) scene material `mat` on the
        device; returns (wi, f[n, 4], pdf[n])."""
        wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
        n = len(wo)
        sample = wi is None
        wi_a = np.zeros((n, 3), np.float32) if sample else np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
        lam_a = None if lam is None else np.ascontiguousarray(lam, np.float32).reshape(n, 4)
        wi_out = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 4), np.float32)
        pdf = np.zeros(n, np.float32)
        self.ctx._check(lib().vs_scene_bsdf(self._h, mat, n, wo.ctypes.data, wi_a.ctypes.data,
                                            None if lam_a is None else lam_a.ctypes.data, 1 if sample else 0, frame,
                                            wi_out.ctypes.data, f.ctypes.data, pdf.ctypes.data))
        return (wi_out if sample else wi_a), f, pdf

    def update_camera(self, c2w: np.ndarray, fov_y: float):
        """Move the camera: `c2w` a 4x4 camera-to-world matrix (acting on
        column vectors), `fov_y` in degrees (the SDK's update_camera)."""
        m = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(4, 4).T)
        self.ctx._check(lib().vs_scene_update_camera(self._h, m.ctypes.data, float(fov_y)))

    def restir_state(self) -> tuple[np.ndarray, np.ndarray]:
        """ReSTIR DI buffers after the last frame: reservoirs float32[4, 3, H*W, 4]
        and surfaces float32[2, 3, H*W, 4] (integer fields as bit patterns)."""
        info = self.info()
        n = info["width"] * info["height"]
        rsv = np.zeros((4, 3, n, 4), np.float32)
        surf = np.zeros((2, 3, n, 4), np.float32)
        self.ctx._check(lib().vs_scene_restir_state(self._h, rsv.ctypes.data, surf.ctypes.data))
        return rsv, surf

    def restir_gi_state(self) -> tuple[np.ndarray, np.ndarray]:
        """ReSTIR GI buffers after the last frame: reservoirs float32[3, 4, H*W, 4]
        (weight_sum C W age | position | normal | Lo) and the indirect
        radiance float32[H*W, 4]."""
        info = self.info()
        n = info["width"] * info["height"]
        rsv = np.zeros((3, 4, n, 4), np.float32)
        rad = np.zeros((n, 4), np.float32)
        self.ctx._check(lib().vs_scene_restir_gi_state(self._h, rsv.ctypes.data, rad.ctypes.data))
        return rsv, rad

    def set_adaptive(self, on: bool = True, threshold: float | None = None, min_samples: int | None = None):
        """Switch adaptive sampling of the pt integrator on or off; threshold and
        minimum sample number default to the scene's "adaptive" block."""
        ex = self.export()["render"]
        t = float(ex[10]) if threshold is None else float(threshold)
        m = int(ex[11]) if min_samples is None else int(min_samples)
        self.ctx._check(lib().vs_scene_set_adaptive(self._h, int(on), t, m))

    def variance_stats(self) -> np.ndarray:
        """Per-pixel VarianceStats float32[H*W, 4] (avg, var, N bits, 0)."""
        info = self.info()
        out = np.zeros((info["width"] * info["height"], 4), np.float32)
        self.ctx._check(lib().vs_scene_variance_stats(self._h, out.ctypes.data))
        return out

    def set_denoiser(self, on: bool = True, reproject: bool = True, filter_moments: bool = True):
        """Switch the rt integrator's SVGF denoiser (and its reprojection and
        moment-filter passes) on or off, as the reference's GUI check boxes do."""
        self.ctx._check(lib().vs_scene_set_denoiser(self._h, int(on), int(reproject), int(filter_moments)))

    def svgf_state(self) -> dict:
        """G-buffer and SVGF buffers after the last denoised frame: gbuffer
        float32[2, 2, H*W, 4], albedo / emission [H*W, 4], records
        [3, 2, H*W, 4], motion [H*W, 4]."""
        info = self.info()
        n = info["width"] * info["height"]
        out = {"gbuffer": np.zeros((2, 2, n, 4), np.float32), "albedo": np.zeros((n, 4), np.float32),
               "emission": np.zeros((n, 4), np.float32), "records": np.zeros((3, 2, n, 4), np.float32),
               "motion": np.zeros((n, 4), np.float32)}
        self.ctx._check(lib().vs_scene_svgf_state(self._h, out["gbuffer"].ctypes.data, out["albedo"].ctypes.data,
                                                  out["emission"].ctypes.data, out["records"].ctypes.data,
                                                  out["motion"].ctypes.data))
        return out

    def set_kernel(self, mode: str = "megakernel", chunk_frames: int = 8, trace_block: int = 128):
        """Kernel form of render(): "megakernel" or "wavefront"."""
        m = {"megakernel": 0, "wavefront": 1}[mode]
        self.ctx._check(lib().vs_scene_set_kernel(self._h, m, chunk_frames, trace_block))

    def render(self, spp: int, fused: bool = True) -> dict:
        self.ctx._check(lib().vs_render(self._h, spp, 1 if fused else 0))
        return self.stats()

    def set_counting(self, on: bool):
        self.ctx._check(lib().vs_set_counting(self._h, 1 if on else 0))

    def stats(self) -> dict:
        s = RenderStats()
        self.ctx._check(lib().vs_last_render_stats(self._h, C.byref(s)))
        out = {}
        for f, _ in RenderStats._fields_:
            v = getattr(s, f)
            out[f] = list(v) if f.startswith("stage_") else v
        return out

    def reset_film(self):
        self.ctx._check(lib().vs_reset_film(self._h))

    def film(self, which: int = 0) -> np.ndarray:
        info = self.info()
        out = np.zeros((info["height"], info["width"], 4), dtype=np.float32)
        self.ctx._check(lib().vs_film_download(self._h, which, out.ctypes.data))
        return out

    def final_picture(self, gamma: bool = True) -> np.ndarray:
        """The picture a save writes: float32[H, W, 4]."""
        info = self.info()
        out = np.zeros((info["height"], info["width"], 4), dtype=np.float32)
        self.ctx._check(lib().vs_film_final_picture(self._h, 1 if gamma else 0, out.ctypes.data))
        return out

    def save(self, path: str | os.PathLike | None = None):
        """Write the final picture; None writes the scene's output.fn."""
        self.ctx._check(lib().vs_film_save(self._h, None if path is None else str(path).encode()))


def _server_check(st: int):
    if st != VS_OK:
        raise VisionError(st, lib().vs_server_last_error().decode())


class Server:
    """Batched serving of path-tracing jobs over resident scenes
    (vs_server_*): continuous batching into persistent kernels, admission by
    an HBM budget, captured step graphs."""

    _CFG = ("hbm_budget", "max_jobs", "step_ms", "tile_w", "waves_per_cu", "use_graphs", "count_rays",
            "min_waves_per_eu", "lane_min_samples")

    def __init__(self, ctx: "Context", hbm_budget: int = 0, max_jobs: int = 0, step_ms: float = 0.0,
                 tile_w: int = 0, waves_per_cu: int = 0, use_graphs: bool = True, count_rays: bool = False,
                 min_waves_per_eu: int = 0, lane_min_samples: int = 0):
        self.ctx = ctx
        self._h = _P()
        cfg = ServerConfig(hbm_budget, max_jobs, step_ms, tile_w, waves_per_cu, 1 if use_graphs else 0,
                           1 if count_rays else 0, min_waves_per_eu, lane_min_samples)
        _server_check(lib().vs_server_create(ctx._h, C.byref(cfg), C.byref(self._h)))
        self.scenes: list = []

    def add_scene(self, scene: "Scene") -> int:
        i = C.c_uint32()
        _server_check(lib().vs_server_add_scene(self._h, scene._h, C.byref(i)))
        self.scenes.append(scene)
        return i.value

    def submit(self, scene: int, width: int, height: int, spp: int, max_depth: int = 0, spectral: bool = False,
               spectrum_dim: int = 0) -> int:
        d = JobDesc(scene, width, height, spp, max_depth, 1 if spectral else 0, spectrum_dim, 0)
        jid = C.c_uint64()
        _server_check(lib().vs_server_submit(self._h, C.byref(d), C.byref(jid)))
        return jid.value

    def wait(self, job: int, timeout_ms: float = -1.0) -> bool:
        done = C.c_int32()
        _server_check(lib().vs_server_wait(self._h, job, timeout_ms, C.byref(done)))
        return bool(done.value)

    def info(self, job: int) -> dict:
        i = JobInfo()
        _server_check(lib().vs_server_job_info(self._h, job, C.byref(i)))
        return {"state": JOB_STATES[i.state], "frames_done": i.frames_done, "samples": i.samples,
                "queue_ms": i.queue_ms, "latency_ms": i.latency_ms,
                "route": "wavefront" if i.route == 1 else "megakernel"}

    def set_route(self, scene: int, spectral: bool, wavefront: bool):
        _server_check(lib().vs_server_set_route(self._h, scene, 1 if spectral else 0, 1 if wavefront else 0))

    def route(self, scene: int, spectral: bool) -> str:
        r = C.c_uint32()
        _server_check(lib().vs_server_get_route(self._h, scene, 1 if spectral else 0, C.byref(r)))
        return "wavefront" if r.value == 1 else "megakernel"

    def result(self, job: int, which: int = 0, shape: Optional[tuple] = None) -> np.ndarray:
        if shape is None:
            raise ValueError("shape=(height, width) is required")
        out = np.zeros((shape[0], shape[1], 4), np.float32)
        _server_check(lib().vs_server_result(self._h, job, which, out.ctypes.data))
        return out

    def release(self, job: int):
        _server_check(lib().vs_server_release(self._h, job))

    def drain(self, timeout_ms: float = -1.0):
        _server_check(lib().vs_server_drain(self._h, timeout_ms))

    def stats(self) -> dict:
        s = ServerStats()
        _server_check(lib().vs_server_stats_get(self._h, C.byref(s)))
        out = {f: getattr(s, f) for f, _ in ServerStats._fields_ if f != "pad"}
        return {f: (list(v) if not isinstance(v, (int, float)) else v) for f, v in out.items()}

    def reset_stats(self):
        _server_check(lib().vs_server_reset_stats(self._h))

    def config(self) -> dict:
        c = ServerConfig()
        _server_check(lib().vs_server_get_config(self._h, C.byref(c)))
        return {f: getattr(c, f) for f in self._CFG}

    def set_config(self, **kw):
        c = ServerConfig()
        _server_check(lib().vs_server_get_config(self._h, C.byref(c)))
        for k, v in kw.items():
            if k not in self._CFG:
                raise KeyError(k)
            setattr(c, k, int(v) if k != "step_ms" else float(v))
        _server_check(lib().vs_server_set_config(self._h, C.byref(c)))

    def autotune(self, probe: list) -> dict:
        """Pick tile width and persistent waves per CU on the probe jobs
        (dicts of submit()'s arguments); returns the trials and the choice."""
        arr = (JobDesc * len(probe))()
        for i, j in enumerate(probe):
            arr[i] = JobDesc(j["scene"], j["width"], j["height"], j["spp"], j.get("max_depth", 0),
                             1 if j.get("spectral") else 0, j.get("spectrum_dim", 0), 0)
        r = AutotuneResult()
        _server_check(lib().vs_server_autotune(self._h, arr, len(probe), C.byref(r)))
        return {"tile_w": r.tile_w, "waves_per_cu": r.waves_per_cu, "min_waves_per_eu": r.min_waves_per_eu,
                "msamples_per_s": r.msamples_per_s,
                "trials": [{"tile_w": r.trial_tile_w[i], "waves_per_cu": r.trial_waves_per_cu[i],
                            "min_waves_per_eu": r.trial_min_waves_per_eu[i],
                            "msamples_per_s": r.trial_msamples_per_s[i]} for i in range(r.trials)],
                "routes": [{"scene": r.route_scene[i], "spectral": bool(r.route_spectral[i]),
                            "route": "wavefront" if r.route_choice[i] else "megakernel",
                            "megakernel_ms": r.route_megakernel_ms[i], "wavefront_ms": r.route_wavefront_ms[i]}
                           for i in range(r.routes)]}

    def close(self):
        if self._h:
            lib().vs_server_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _nrc_check(st: int):
    if st != VS_OK:
        raise VisionError(st, lib().vs_nrc_last_error().decode())


class RadianceCache:
    """Online-trained radiance cache of an sRGB scene (vs_nrc_*)."""

    def __init__(self, scene: "Scene", **cfg):
        c = NrcConfig()
        lib().vs_nrc_default_config(C.byref(c))
        for k, v in cfg.items():
            if k not in dict(NrcConfig._fields_):
                raise KeyError(k)
            setattr(c, k, float(v) if k == "learning_rate" else int(v))
        self.scene = scene
        self._h = _P()
        _nrc_check(lib().vs_nrc_create(scene._h, C.byref(c), C.byref(self._h)))

    def render(self, frames: int, spp: int, terminate: bool = True, train: bool = True):
        flags = (NRC_TERMINATE if terminate else 0) | (NRC_TRAIN if train else 0)
        _nrc_check(lib().vs_nrc_render(self._h, frames, spp, flags))
        return self.stats()

    def stats(self) -> dict:
        s = NrcStats()
        _nrc_check(lib().vs_nrc_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in NrcStats._fields_}

    def levels(self) -> dict:
        res = np.zeros(16, np.uint32)
        mask = C.c_uint32()
        lo = np.zeros(3, np.float32)
        inv = np.zeros(3, np.float32)
        _nrc_check(lib().vs_nrc_levels(self._h, res.ctypes.data, C.byref(mask), lo.ctypes.data, inv.ctypes.data))
        return {"res": res, "dense_mask": mask.value, "lo": lo, "inv_ext": inv}

    def params(self, grid: bool = True):
        mlp = np.zeros(NRC_MLP_PARAMS, np.float32)
        g = np.zeros(NRC_GRID_PARAMS, np.float32) if grid else None
        _nrc_check(lib().vs_nrc_get_params(self._h, mlp.ctypes.data, None if g is None else g.ctypes.data))
        return mlp, g

    def set_params(self, mlp: np.ndarray, grid: Optional[np.ndarray] = None):
        mlp = np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        _nrc_check(lib().vs_nrc_set_params(self._h, mlp.ctypes.data, None if g is None else g.ctypes.data))

    def forward(self, pos: np.ndarray) -> np.ndarray:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        out = np.zeros((len(pos), 3), np.float32)
        _nrc_check(lib().vs_nrc_forward(self._h, pos.ctypes.data, len(pos), out.ctypes.data))
        return out

    def gradients(self, samples: np.ndarray, grid: bool = True):
        sm = np.ascontiguousarray(samples, np.float32).reshape(-1, 6)
        loss = C.c_float()
        gm = np.zeros(NRC_MLP_PARAMS, np.float32)
        gg = np.zeros(NRC_GRID_PARAMS, np.float32) if grid else None
        _nrc_check(lib().vs_nrc_gradients(self._h, sm.ctypes.data, len(sm), C.byref(loss), gm.ctypes.data,
                                          None if gg is None else gg.ctypes.data))
        return loss.value, gm, gg

    def train(self, samples: np.ndarray, steps: int = 1) -> float:
        sm = np.ascontiguousarray(samples, np.float32).reshape(-1, 6)
        loss = C.c_float()
        _nrc_check(lib().vs_nrc_train(self._h, sm.ctypes.data, len(sm), steps, C.byref(loss)))
        return loss.value

    def close(self):
        if self._h:
            lib().vs_nrc_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
