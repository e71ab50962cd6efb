"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product (vision_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
# numeric tables (data files, not code): energy-compensation LUTs, CIE 1931
# and D65 at 1 nm, and the generated sRGB -> sigmoid coefficient table
DATA = HERE.parent / "vision_amd" / "data"
_lib = None
_data: dict = {}

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def build() -> Path:
    subprocess.run(["make", "-B", "-C", str(HERE)], check=True, capture_output=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.or_tea.restype = C.c_uint32
        h.or_tea.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]
        h.or_tea_v1.restype = C.c_uint32
        h.or_tea_v1.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]
        h.or_sampler_stream.restype = None
        h.or_sampler_stream.argtypes = [C.c_uint32] * 5 + [_f32p, C.c_void_p]
        h.or_lcg.argtypes = [C.c_uint32, C.c_uint32, _u32p, _f32p]
        h.or_cosine_hemisphere.argtypes = [_f32p, C.c_uint32, _f32p]
        h.or_square_to_triangle.argtypes = [_f32p, C.c_uint32, _f32p]
        h.or_sample_tent.argtypes = [_f32p, C.c_uint32, C.c_float, _f32p]
        h.or_alias_build.argtypes = [_f32p, C.c_uint32, _f32p, _u32p, C.POINTER(C.c_float)]
        h.or_alias_sample.argtypes = [_f32p, C.c_uint32, _f32p, C.c_uint32, _u32p, C.c_void_p, C.c_void_p]
        h.or_offset_ray_origin.argtypes = [_f32p, _f32p, C.c_uint32, _f32p]
        h.or_filter_eval.argtypes = [_f32p, _f32p, C.c_uint32, _f32p]
        h.or_film_output.argtypes = [_f32p, C.c_uint32, C.c_float, C.c_uint32, _f32p]
        h.or_elementary.argtypes = [C.c_int, _f32p, _f32p, C.c_uint32, _f32p]
        h.or_camera_rays.argtypes = [_f32p] + [C.c_uint32] * 4 + [_f32p, _f32p, _f32p]
        _f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        h.or_nrc_levels.argtypes = [C.c_uint32, C.c_uint32, _u32p, C.POINTER(C.c_uint32)]
        h.or_nrc_encode.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, C.c_int32, _f64p]
        h.or_nrc_forward.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, C.c_uint32, C.c_int32, _f64p]
        h.or_nrc_backward.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, C.c_uint32, C.c_int32, C.POINTER(C.c_double),
                                      _f64p, C.c_void_p]
        h.or_nrc_adam.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_uint32, C.c_float, C.c_float, C.c_float]
        h.or_nrc_param_count.restype = C.c_uint32
        h.or_final_picture.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int32, _f32p]
        h.or_filter_table.argtypes = [_f32p, _f32p]
        h.or_filter_sample.argtypes = [_f32p, _f32p, C.c_uint32, _f32p, _f32p]
        h.or_eval_diffuse.argtypes = [C.c_uint32, _f32p, C.c_float, _f32p, _f32p, C.c_uint32, _f32p, _f32p]
        h.or_trace.argtypes = [_f32p, _u32p, C.c_uint32, C.c_uint32, _f32p, C.c_uint32, C.c_int, C.c_int,
                               _u32p, _f32p, _f32p]
        h.or_render.restype = C.c_uint64
        h.or_render.argtypes = ([_f32p, _f32p, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p, _f32p, _f32p] +
                                [C.c_uint32] * 5 + [C.c_uint32] * 6 + [C.c_int, C.c_uint32, C.c_void_p, _f32p,
                                                                       C.c_void_p])
        h.or_restir_frame.argtypes = ([_f32p, _f32p, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p, _f32p, _f32p] +
                                      [C.c_uint32] * 5 + [C.c_void_p, _f32p, _f32p, _f32p, C.c_uint32, C.c_uint32,
                                                          _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                          C.c_void_p])
        h.or_gbuffer_gradient.argtypes = [_f32p, C.c_uint32, C.c_uint32]
        h.or_svgf_frame.argtypes = [_f32p, C.c_uint32, C.c_uint32, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                    _f32p, _f32p]
        h.or_variance_stats_add.argtypes = [_f32p, _f32p, C.c_uint32]
        h.or_reservoir_stream.restype = C.c_int32
        h.or_reservoir_stream.argtypes = [_f32p, _f32p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        h.or_prev_raster.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, _f32p]
        h.or_cie_tables.argtypes = [_f32p, _f32p, C.POINTER(C.c_float)]
        h.or_spd_eval.argtypes = [_f32p, C.c_float, _f32p, C.c_uint32, _f32p]
        h.or_wavelengths.argtypes = [_f32p, C.c_uint32, _f32p, _f32p]
        h.or_wavelength_pdf.argtypes = [_f32p, C.c_uint32, _f32p]
        h.or_sigmoid.argtypes = [_f32p, C.c_uint32, _f32p, C.c_uint32, _f32p]
        h.or_rgb_to_spectrum.argtypes = [_f32p, _f32p, C.c_uint32, C.c_int, _f32p]
        h.or_fresnel.argtypes = [C.c_int, _f32p, C.c_uint32, C.c_float, C.c_float, _f32p]
        h.or_ggx.argtypes = [C.c_float, C.c_float, _f32p, _f32p, C.c_uint32, _f32p, _f32p, _f32p]
        h.or_ggx_sample.argtypes = [C.c_float, C.c_float, _f32p, _f32p, C.c_uint32, _f32p]
        h.or_bsdf_eval.argtypes = [C.c_void_p, _u32p, _f32p, _f32p, C.c_void_p, C.c_uint32, _f32p, _f32p]
        h.or_bsdf_sample.argtypes = [C.c_void_p, _u32p, _f32p, C.c_void_p, C.c_uint32, C.c_uint32, _f32p, _f32p,
                                     _f32p]
        _lib = h
    return _lib


def tea(v0: int, v1: int, key=None, rounds: int = 4) -> int:
    k = None if key is None else (C.c_uint32 * 4)(*key)
    return lib().or_tea(v0, v1, k, rounds)


def tea_block(v0: int, v1: int, key, rounds: int = 32) -> tuple[int, int]:
    k = (C.c_uint32 * 4)(*key)
    return lib().or_tea(v0, v1, k, rounds), lib().or_tea_v1(v0, v1, k, rounds)


def sampler_stream(px, py, index, dim, count):
    out = np.zeros(count, np.float32)
    states = np.zeros(count + 1, np.uint32)
    lib().or_sampler_stream(px, py, index, dim, count, out, states.ctypes.data)
    return out, states


def lcg(state: int, count: int):
    states = np.zeros(count, np.uint32)
    values = np.zeros(count, np.float32)
    lib().or_lcg(state, count, states, values)
    return states, values


def cosine_hemisphere(u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((len(u), 3), np.float32)
    lib().or_cosine_hemisphere(u, len(u), out)
    return out


def square_to_triangle(u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((len(u), 2), np.float32)
    lib().or_square_to_triangle(u, len(u), out)
    return out


def sample_tent(u: np.ndarray, r: float) -> np.ndarray:
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros(len(u), np.float32)
    lib().or_sample_tent(u, len(u), r, out)
    return out


def alias_build(w: np.ndarray):
    w = np.ascontiguousarray(w, np.float32)
    prob = np.zeros(len(w), np.float32)
    alias = np.zeros(len(w), np.uint32)
    integral = C.c_float()
    lib().or_alias_build(w, len(w), prob, alias, C.byref(integral))
    return prob, alias, integral.value


def alias_sample(w: np.ndarray, u: np.ndarray):
    w = np.ascontiguousarray(w, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    idx = np.zeros(len(u), np.uint32)
    urem = np.zeros(len(u), np.float32)
    pmf = np.zeros(len(u), np.float32)
    lib().or_alias_sample(w, len(w), u, len(u), idx, urem.ctypes.data, pmf.ctypes.data)
    return idx, urem, pmf


FILTER_KINDS = {"box": 0, "triangle": 1, "gaussian": 2, "mitchell": 3, "sinc": 4}


def filter_camera(kind: str, radius: float, sigma=1.0, b=1 / 3, c=1 / 3, tau=3.0) -> np.ndarray:
    """A 32-float camera record carrying only the pixel-filter fields."""
    cam = np.zeros(32, np.float32)
    k = FILTER_KINDS[kind]
    cam[21] = k if k <= 1 else 2
    cam[22] = radius
    cam[24:29] = [k, sigma, b, c, tau]
    return cam


def filter_eval(cam: np.ndarray, xy: np.ndarray) -> np.ndarray:
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    out = np.zeros(len(xy), np.float32)
    lib().or_filter_eval(np.ascontiguousarray(cam, np.float32), xy, len(xy), out)
    return out


TONE_MAPPERS = {"linear": 0, "aces": 1, "reinhard": 2}


def film_output(rt: np.ndarray, exposure: float = 1.0, tone_mapper: str = "linear") -> np.ndarray:
    """The collector's output buffer from its running-mean film [..., 4]."""
    a = np.ascontiguousarray(rt, np.float32)
    out = np.zeros_like(a)
    lib().or_film_output(a.reshape(-1), a.size // 4, exposure, TONE_MAPPERS[tone_mapper], out.reshape(-1))
    return out


def final_picture(output: np.ndarray, tone_mapper: str = "linear", gamma: bool = True) -> np.ndarray:
    """The picture a save writes, from the output buffer [..., 4]."""
    a = np.ascontiguousarray(output, np.float32)
    out = np.zeros_like(a)
    lib().or_final_picture(a.reshape(-1), a.size // 4, TONE_MAPPERS[tone_mapper], 1 if gamma else 0,
                           out.reshape(-1))
    return out


def filter_table(cam: np.ndarray) -> np.ndarray:
    lut = np.zeros(400, np.float32)
    lib().or_filter_table(np.ascontiguousarray(cam, np.float32), lut)
    return lut.reshape(20, 20)


def filter_sample(cam: np.ndarray, u: np.ndarray):
    """Filter importance sampling: film-plane offsets (n, 2) and weights (n,)."""
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    p = np.zeros_like(u)
    w = np.zeros(len(u), np.float32)
    lib().or_filter_sample(np.ascontiguousarray(cam, np.float32), u, len(u), p, w)
    return p, w


def offset_ray_origin(p: np.ndarray, n: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(n, np.float32).reshape(-1, 3)
    out = np.zeros_like(p)
    lib().or_offset_ray_origin(p, n, len(p), out)
    return out


def eval_diffuse(type_: int, color, sigma: float, wo: np.ndarray, wi: np.ndarray):
    color = np.ascontiguousarray(color, np.float32)
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    f = np.zeros_like(wo)
    pdf = np.zeros(len(wo), np.float32)
    lib().or_eval_diffuse(type_, color, sigma, wo, wi, len(wo), f, pdf)
    return f, pdf


def trace(positions: np.ndarray, triangles: np.ndarray, rays: np.ndarray, any_hit: bool = False,
          bvh: bool = True):
    """rays: float32[n, 8] (org, t_min, dir, t_max).  Returns (tri, t, bary)."""
    P = np.ascontiguousarray(positions, np.float32).reshape(-1)
    T = np.ascontiguousarray(triangles, np.uint32).reshape(-1)
    R = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(R)
    tri = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    bary = np.zeros((n, 2), np.float32)
    lib().or_trace(P, T, len(P) // 3, len(T) // 3, R, n, int(any_hit), 1 if bvh else 0, tri, t, bary)
    return tri, t, bary


def render(scene: dict, frame_start: int = 0, spp: int = 1, window=None, bvh: bool = True, threads: int = 0,
           film: np.ndarray | None = None, vstats: np.ndarray | None = None) -> np.ndarray:
    """Render the exported scene (vision_amd.Scene.export()) on the CPU.
    Returns the running-mean film float32[H, W, 4] (updated in place if given).
    With adaptive sampling on in the export, `vstats` float32[H*W, 4] holds
    the per-pixel statistics (updated in place; zero them at frame 0)."""
    spec = Spectral(scene) if "spd" in scene else None
    cam = np.ascontiguousarray(scene["camera"], np.float32)
    W, H = int(cam[19]), int(cam[20])
    if film is None:
        film = np.zeros((H, W, 4), np.float32)
    x0, y0, x1, y1 = window if window is not None else (0, 0, W, H)
    a = {k: np.ascontiguousarray(v) for k, v in scene.items() if isinstance(v, np.ndarray)}
    lib().or_render(a["positions"].reshape(-1), a["normals"].reshape(-1), a["uvs"].reshape(-1),
                    a["triangles"].reshape(-1), a["tri_instance"], a["instances"].reshape(-1),
                    a["materials"].reshape(-1), a["lights"].reshape(-1), cam, a["render"],
                    len(a["positions"]), len(a["triangles"]), len(a["instances"]), len(a["materials"]),
                    len(a["lights"]), x0, y0, x1, y1, frame_start, spp, 1 if bvh else 0, threads,
                    None if spec is None else spec.ptr, film, None if vstats is None else vstats.ctypes.data)
    return film


def gbuffer_gradient(gbuf: np.ndarray, W: int, H: int) -> np.ndarray:
    """FrameBuffer::compute_gradient over a PixelGeometry buffer float32[H*W, 8]
    (normal.xyz fwidth | p_film.xy depth_gradient depth); returns a copy
    with fwidth and the depth gradient filled."""
    g = np.ascontiguousarray(gbuf, np.float32).reshape(-1).copy()
    lib().or_gbuffer_gradient(g, W, H)
    return g.reshape(H * W, 8)


SVGF_DEFAULTS = dict(N=3, alpha=0.05, moments_alpha=0.2, history_limit=32, moments_filter_radius=3, sigma_rt=10.0,
                     sigma_normal=30.0, reproject=True, filter_moments=True)


def svgf_params(**kw) -> np.ndarray:
    p = dict(SVGF_DEFAULTS, **kw)
    return np.array([p["N"], p["alpha"], p["moments_alpha"], p["history_limit"], p["moments_filter_radius"],
                     p["sigma_rt"], p["sigma_normal"], float(p["reproject"]), float(p["filter_moments"])], np.float32)


def svgf_frame(params: np.ndarray, W: int, H: int, gcur, gprev, motion, rad, albedo, emission, prev_rec):
    """One SVGF frame (or_svgf_frame): G-buffers float32[H*W, 8], motion
    [H*W, 2], radiance / albedo / emission [H*W, 4], the previous frame's
    records [H*W, 8].  Returns (records [H*W, 8], picture [H*W, 4])."""
    f = lambda a, n: np.ascontiguousarray(a, np.float32).reshape(H * W, n)
    cur = np.zeros((H * W, 8), np.float32)
    out = np.zeros((H * W, 4), np.float32)
    lib().or_svgf_frame(np.ascontiguousarray(params, np.float32), W, H, f(gcur, 8), f(gprev, 8), f(motion, 2),
                        f(rad, 4), f(albedo, 4), f(emission, 4), f(prev_rec, 8), cur, out)
    return cur, out


def variance_stats_add(stats: np.ndarray, x: np.ndarray) -> np.ndarray:
    """VarianceStats::add on per-pixel statistics float32[n, 4] (avg, var, N
    bits, 0), in place."""
    lib().or_variance_stats_add(stats, np.ascontiguousarray(x, np.float32), len(x))
    return stats


def reservoir_stream(w: np.ndarray, u: np.ndarray):
    """Stream weights through one reservoir: (kept index or -1, weight_sum, C)."""
    w = np.ascontiguousarray(w, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    ws, c = C.c_float(), C.c_float()
    k = lib().or_reservoir_stream(w, u, len(w), C.byref(ws), C.byref(c))
    return k, ws.value, c.value


def prev_raster(cam: np.ndarray, w2c: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Raster coordinates [n, 2] of world points in the camera (cam record, w2c 3x4)."""
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    out = np.zeros((len(p), 2), np.float32)
    lib().or_prev_raster(np.ascontiguousarray(cam, np.float32), np.ascontiguousarray(w2c, np.float32), p, len(p), out)
    return out


class Restir:
    """ReSTIR DI frames on the CPU (or_restir_frame) with the state carried
    between frames: reservoirs float32[4, 3, P, 4], surfaces [2, 3, P, 4],
    surface extensions [2, 2, P, 4] and the running-mean film [H, W, 4].
    The camera of the previous frame is the one of the last rendered scene
    export; a moved camera is passed as a new export."""

    def __init__(self, scene: dict, threads: int = 0):
        cam = scene["camera"]
        self.W, self.H = int(cam[19]), int(cam[20])
        P = self.W * self.H
        self.rsv = np.zeros((4, 3, P, 4), np.float32)
        self.rsv[:, 1, :, 0] = np.frombuffer(np.uint32(0xFFFFFFFF).tobytes(), np.float32)[0]
        self.surf = np.zeros((2, 3, P, 4), np.float32)
        self.surf[:, 0, :, 0] = np.frombuffer(np.uint32(0xFFFFFFFF).tobytes(), np.float32)[0]
        self.surf[:, 0, :, 3] = np.frombuffer(np.uint32(1).tobytes(), np.float32)[0]
        self.ext = np.zeros((2, 2, P, 4), np.float32)
        self.ext[:, 0, :, :3] = 1.0
        self.film = np.zeros((self.H, self.W, 4), np.float32)
        # ReSTIR GI: reservoirs [3][4][P], samples [4][P], DI's per-pixel
        # record [3][P] (HitBSDF, motion), indirect radiance [P]
        self.gi_rsv = np.zeros((3, 4, P, 4), np.float32)
        self.gi_smp = np.zeros((4, P, 4), np.float32)
        self.gi_aux = np.zeros((3, P, 4), np.float32)
        self.gi_rad = np.zeros((P, 4), np.float32)
        # G-buffer [2][P][8] (PixelGeometry, depth -1 when empty), albedo /
        # emission [P][4], SVGF records [2][P][8], the denoised picture
        self.geo = np.zeros((2, P, 8), np.float32)
        self.geo[:, :, 7] = -1.0
        self.alb = np.zeros((P, 4), np.float32)
        self.emi = np.zeros((P, 4), np.float32)
        self.svgf_rec = np.zeros((2, P, 8), np.float32)
        self.picture = np.zeros((P, 4), np.float32)
        self.frame = 0
        self.threads = threads
        self.scene = scene
        self.prev_w2c = np.ascontiguousarray(scene["restir"][20:32], np.float32)
        self.prev_pos = np.array([cam[3], cam[7], cam[11]], np.float32)

    def render(self, frames: int = 1, scene: dict | None = None) -> np.ndarray:
        if scene is not None:
            self.scene = scene
        sc = self.scene
        spec = Spectral(sc)
        cam = np.ascontiguousarray(sc["camera"], np.float32)
        a = {k: np.ascontiguousarray(v) for k, v in sc.items() if isinstance(v, np.ndarray)}
        svgf = float(sc["restir"][48]) != 0.0 if len(sc["restir"]) > 48 else False
        P = self.W * self.H
        for _ in range(frames):
            cur, prev = (self.frame + 1) & 1, self.frame & 1
            gblock = np.zeros(16 * P, np.float32) if svgf else None
            lib().or_restir_frame(a["positions"].reshape(-1), a["normals"].reshape(-1), a["uvs"].reshape(-1),
                                  a["triangles"].reshape(-1), a["tri_instance"], a["instances"].reshape(-1),
                                  a["materials"].reshape(-1), a["lights"].reshape(-1), cam, a["render"],
                                  len(a["positions"]), len(a["triangles"]), len(a["instances"]),
                                  len(a["materials"]), len(a["lights"]), spec.ptr,
                                  np.ascontiguousarray(a["restir"], np.float32), self.prev_w2c, self.prev_pos,
                                  self.frame, self.threads, self.rsv, self.surf, self.ext, self.film, self.gi_rsv,
                                  self.gi_smp, self.gi_aux, self.gi_rad,
                                  None if gblock is None else gblock.ctypes.data)
            if svgf:
                # RealTimeIntegrator::render: the denoiser after the combine
                # (rt.cpp:157-158)
                self.geo[cur] = gblock[:8 * P].reshape(P, 8)
                self.alb = gblock[8 * P:12 * P].reshape(P, 4).copy()
                self.emi = gblock[12 * P:].reshape(P, 4).copy()
                rs = np.asarray(sc["restir"], np.float32)
                params = rs[49:58]
                rec, pic = svgf_frame(params, self.W, self.H, self.geo[cur], self.geo[prev],
                                      self.gi_aux[2, :, :2], self.film.reshape(P, 4), self.alb, self.emi,
                                      self.svgf_rec[prev])
                self.svgf_rec[cur] = rec
                self.picture = pic
            self.frame += 1
            self.prev_w2c = np.ascontiguousarray(a["restir"][20:32], np.float32)
            self.prev_pos = np.array([cam[3], cam[7], cam[11]], np.float32)
        return self.film


# ---------------------------------------------------------------- spectral inputs
class OrSpectral(C.Structure):
    """Mirror of `or_spectral` in oracle.cpp."""
    _fields_ = [("rgb", C.c_uint32), ("ns", C.c_uint32), ("spd", C.c_void_p), ("cie_y_integral", C.c_float),
                ("lut_diel", C.c_void_p), ("lut_diel_inv", C.c_void_p), ("lut_spec", C.c_void_p),
                ("lut_coat", C.c_void_p), ("lut_pure", C.c_void_p), ("lut_sheen", C.c_void_p),
                ("rgb2spec", C.c_void_p), ("env_light", C.c_uint32), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("env_rgb", C.c_void_p), ("env_w2o", C.c_float * 9), ("env_o2w", C.c_float * 9),
                ("env_scale", C.c_float), ("world_radius", C.c_float), ("tex", C.c_void_p),
                ("tex_info", C.c_void_p), ("ntex", C.c_uint32)]


def data(name: str, count: int | None = None) -> np.ndarray:
    """A float32 data table from vision_amd/data (cached)."""
    if name not in _data:
        arr = np.fromfile(DATA / name, dtype="<f4")
        if count is not None and arr.size != count:
            raise ValueError(f"{name}: {arr.size} floats, expected {count}")
        _data[name] = np.ascontiguousarray(arr, np.float32)
    return _data[name]


LUTS = {"lut_diel": ("lut_dielectric.f32", 65536), "lut_diel_inv": ("lut_dielectric_inv.f32", 65536),
        "lut_spec": ("lut_specular.f32", 32768), "lut_coat": ("lut_coat.f32", 32768),
        "lut_pure": ("lut_pure_reflection.f32", 1024), "lut_sheen": ("lut_sheen_approx.f32", 4096)}


def cie_tables() -> tuple[np.ndarray, float]:
    """(CIE X, Y, Z, D65 downsampled to 4 x 94, CIE Y integral) from the 1 nm data."""
    raw = data("cie_xyz_d65.f32", 4 * 471)
    out = np.zeros(4 * 94, np.float32)
    yi = C.c_float()
    lib().or_cie_tables(raw, out, C.byref(yi))
    return out, yi.value


def rgb2spec_table() -> np.ndarray:
    return data("rgb2spec_srgb.f32", 3 * 64 * 64 * 64 * 3)


class Spectral:
    """Holds the arrays an OrSpectral points at (keep it alive during calls)."""

    def __init__(self, scene: dict | None = None, rgb: bool = True, ns: int = 3, spd: np.ndarray | None = None):
        if scene is not None:
            rgb = bool(scene["spectrum_rgb"])
            ns = 3 if rgb else int(scene["spectrum_dim"])
            spd = scene["spd"]
        cie, y_integral = cie_tables()
        self.spd = np.ascontiguousarray(spd if spd is not None else cie, np.float32)
        self.luts = {k: data(*v) for k, v in LUTS.items()}
        self.table = rgb2spec_table() if not rgb else np.zeros(1, np.float32)
        s = OrSpectral()
        s.rgb = 1 if rgb else 0
        s.ns = ns
        s.spd = self.spd.ctypes.data
        s.cie_y_integral = y_integral
        for k, v in self.luts.items():
            setattr(s, k, v.ctypes.data)
        s.rgb2spec = self.table.ctypes.data
        s.env_light = 0xFFFFFFFF
        s.env_scale = 1.0
        s.world_radius = 1.0
        if scene is not None:
            s.world_radius = scene["world_radius"]
            if len(scene.get("texture_info", [])):
                self.tex = np.ascontiguousarray(scene["textures"], np.float32)
                self.tex_info = np.ascontiguousarray(scene["texture_info"], np.uint32)
                s.tex = self.tex.ctypes.data
                s.tex_info = self.tex_info.ctypes.data
                s.ntex = len(self.tex_info)
            if scene["env_light"] != 0xFFFFFFFF:
                self.env_rgb = np.ascontiguousarray(scene["env_rgb"], np.float32)
                s.env_light = scene["env_light"]
                s.env_h, s.env_w = self.env_rgb.shape[:2]
                s.env_rgb = self.env_rgb.ctypes.data
                s.env_w2o[:] = [float(x) for x in np.asarray(scene["env_w2o"]).reshape(-1)]
                s.env_o2w[:] = [float(x) for x in np.asarray(scene["env_o2w"]).reshape(-1)]
                s.env_scale = scene["env_scale"]
        self.s = s

    @property
    def ptr(self):
        return C.byref(self.s)


def spd_eval(tab: np.ndarray, interval: float, lam: np.ndarray) -> np.ndarray:
    tab = np.ascontiguousarray(tab, np.float32)
    lam = np.ascontiguousarray(lam, np.float32).reshape(-1)
    out = np.zeros_like(lam)
    lib().or_spd_eval(tab, interval, lam, len(lam), out)
    return out


def wavelengths(u: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    u = np.ascontiguousarray(u, np.float32).reshape(-1)
    lam = np.zeros_like(u)
    pdf = np.zeros_like(u)
    lib().or_wavelengths(u, len(u), lam, pdf)
    return lam, pdf


def wavelength_pdf(lam: np.ndarray) -> np.ndarray:
    lam = np.ascontiguousarray(lam, np.float32).reshape(-1)
    out = np.zeros_like(lam)
    lib().or_wavelength_pdf(lam, len(lam), out)
    return out


def sigmoid(coef: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """Sigmoid spectra float32[ncoef, nlambda] of coefficient rows (c0, c1, c2, s)."""
    coef = np.ascontiguousarray(coef, np.float32).reshape(-1, 4)
    lam = np.ascontiguousarray(lam, np.float32).reshape(-1)
    out = np.zeros((len(coef), len(lam)), np.float32)
    lib().or_sigmoid(coef, len(coef), lam, len(lam), out)
    return out


def rgb_to_spectrum(rgb: np.ndarray, unbound: bool = False) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros((len(rgb), 4), np.float32)
    lib().or_rgb_to_spectrum(rgb2spec_table(), rgb, len(rgb), 1 if unbound else 0, out)
    return out


def fresnel(cos_t: np.ndarray, eta: float, k: float | None = None) -> np.ndarray:
    cos_t = np.ascontiguousarray(cos_t, np.float32).reshape(-1)
    out = np.zeros_like(cos_t)
    lib().or_fresnel(0 if k is not None else 1, cos_t, len(cos_t), eta, 0.0 if k is None else k, out)
    return out


def ggx(ax: float, ay: float, wo: np.ndarray, wh: np.ndarray):
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wh = np.ascontiguousarray(wh, np.float32).reshape(-1, 3)
    n = len(wo)
    D, G1, pdf = (np.zeros(n, np.float32) for _ in range(3))
    lib().or_ggx(ax, ay, wo, wh, n, D, G1, pdf)
    return D, G1, pdf


def ggx_sample(ax: float, ay: float, wo: np.ndarray, u: np.ndarray) -> np.ndarray:
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    out = np.zeros_like(wo)
    lib().or_ggx_sample(ax, ay, wo, u, len(wo), out)
    return out


def material_record(type_: int, params, flags: int = 0, uint_slots: dict | None = None) -> np.ndarray:
    """A 32-word material record: type, flags, 30 float parameters; uint_slots
    maps a parameter index to an integer stored bitwise (SPD offsets)."""
    rec = np.zeros(32, np.uint32)
    rec[0] = type_
    rec[1] = flags
    p = np.zeros(30, np.float32)
    p[:len(params)] = params
    rec[2:] = p.view(np.uint32)
    for k, v in (uint_slots or {}).items():
        rec[2 + k] = v
    return rec


def bsdf_eval(spec: Spectral, rec: np.ndarray, wo: np.ndarray, wi: np.ndarray, lam: np.ndarray | None = None):
    """f * |cos wi| (float32[n, 4]) and pdf of material `rec` in a frame whose
    shading and geometric normal is +z."""
    rec = np.ascontiguousarray(rec, np.uint32)
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    n = len(wo)
    lam_a = None if lam is None else np.ascontiguousarray(lam, np.float32).reshape(n, 4)
    f = np.zeros((n, 4), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().or_bsdf_eval(spec.ptr, rec, wo, wi, None if lam_a is None else lam_a.ctypes.data, n, f, pdf)
    return f, pdf


def bsdf_sample(spec: Spectral, rec: np.ndarray, wo: np.ndarray, frame: int = 0, lam: np.ndarray | None = None):
    """Sampled wi (float32[n, 3]), f * |cos wi| and pdf for each wo."""
    rec = np.ascontiguousarray(rec, np.uint32)
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    n = len(wo)
    lam_a = None if lam is None else np.ascontiguousarray(lam, np.float32).reshape(n, 4)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 4), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().or_bsdf_sample(spec.ptr, rec, wo, None if lam_a is None else lam_a.ctypes.data, n, frame, wi, f, pdf)
    return wi, f, pdf


# ---- radiance cache (oracle/nrc.cpp) ----------------------------------------------

NRC_GRID = 16 * (1 << 19) * 2


class NrcCfg(C.Structure):
    _fields_ = [("res", C.c_uint32 * 16), ("dense_mask", C.c_uint32), ("lo", C.c_float * 3),
                ("inv_ext", C.c_float * 3)]


def nrc_cfg(res, dense_mask, lo, inv_ext) -> NrcCfg:
    c = NrcCfg()
    for i in range(16):
        c.res[i] = int(res[i])
    c.dense_mask = int(dense_mask)
    for i in range(3):
        c.lo[i] = float(lo[i])
        c.inv_ext[i] = float(inv_ext[i])
    return c


def nrc_levels(min_res: int, max_res: int):
    res = np.zeros(16, np.uint32)
    mask = C.c_uint32()
    lib().or_nrc_levels(min_res, max_res, res, C.byref(mask))
    return res, mask.value


def nrc_encode(cfg: NrcCfg, grid: np.ndarray, pos: np.ndarray, round: bool = True) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.zeros((len(pos), 32), np.float64)
    lib().or_nrc_encode(C.byref(cfg), np.ascontiguousarray(grid, np.float32), pos, len(pos), 1 if round else 0, out)
    return out


def nrc_forward(cfg: NrcCfg, mlp: np.ndarray, grid: np.ndarray, pos: np.ndarray, round: bool = True) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.zeros((len(pos), 3), np.float64)
    lib().or_nrc_forward(C.byref(cfg), np.ascontiguousarray(mlp, np.float32), np.ascontiguousarray(grid, np.float32),
                         pos, len(pos), 1 if round else 0, out)
    return out


def nrc_backward(cfg: NrcCfg, mlp: np.ndarray, grid: np.ndarray, samples: np.ndarray, round: bool = True,
                 with_grid: bool = True):
    sm = np.ascontiguousarray(samples, np.float32).reshape(-1, 6)
    loss = C.c_double()
    gm = np.zeros(len(mlp), np.float64)
    gg = np.zeros(NRC_GRID, np.float64) if with_grid else None
    lib().or_nrc_backward(C.byref(cfg), np.ascontiguousarray(mlp, np.float32), np.ascontiguousarray(grid, np.float32),
                          sm, len(sm), 1 if round else 0, C.byref(loss), gm,
                          None if gg is None else gg.ctypes.data)
    return loss.value, gm, gg


def nrc_adam(p, m, v, g, lr: float, t: int):
    c1 = 1.0 / (1.0 - 0.9 ** t)
    c2 = 1.0 / (1.0 - 0.99 ** t)
    lib().or_nrc_adam(p, m, v, g, len(p), lr, c1, c2)


ELEMENTARY = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "log1p": 4, "atanh": 5, "cosh": 6, "acos": 7, "atan2": 8,
              "pow": 9}


def elementary(name: str, x, y=None) -> np.ndarray:
    """The library's shared single-precision elementary function `name` over
    float32 arrays (atan2 takes (y, x) as (x, y) here: elementary("atan2",
    y, x); pow takes (base, exponent))."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    yy = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float32)
    out = np.empty_like(x)
    lib().or_elementary(ELEMENTARY[name], x, yy, x.size, out)
    return out


def camera_rays(cam: np.ndarray, px: int, py: int, frame0: int, n: int):
    """World-space origins, directions (float32[n, 3]) and film positions
    (float32[n, 2]) of pixel (px, py)'s camera rays on frames frame0..+n."""
    cam = np.ascontiguousarray(cam, np.float32).reshape(-1)
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    pf = np.zeros((n, 2), np.float32)
    lib().or_camera_rays(cam, px, py, frame0, n, o, d, pf)
    return o, d, pf


def material_table(materials: np.ndarray, m: int) -> np.ndarray:
    """The records material `m` of a scene export needs, as a table whose
    first record is `m` (the form bsdf_eval / bsdf_sample take): `m` alone,
    or a mix / add followed by its two components with indices 1 and 2."""
    mats = np.ascontiguousarray(materials, np.uint32).reshape(-1, 32)
    rec = mats[m].copy()
    if rec[0] not in (7, 8):
        return rec
    comps = [mats[int(rec[2])], mats[int(rec[3])]]
    rec[2], rec[3] = 1, 2
    return np.concatenate([rec, *comps])
