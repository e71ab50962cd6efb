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
_lib = None

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
        h.or_eval_diffuse.argtypes = [C.c_uint32, _f32p, C.c_float, _f32p, _f32p, C.c_uint32, _f32p, _f32p]
        h.or_trace.argtypes = [_f32p, _u32p, C.c_uint32, C.c_uint32, _f32p, C.c_uint32, C.c_int, C.c_int,
                               _u32p, _f32p, _f32p]
        h.or_render.restype = C.c_uint64
        h.or_render.argtypes = ([_f32p, _f32p, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p, _f32p, _f32p] +
                                [C.c_uint32] * 5 + [C.c_uint32] * 6 + [C.c_int, C.c_uint32, _f32p])
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
           film: np.ndarray | None = None) -> np.ndarray:
    """Render the exported scene (vision_amd.Scene.export()) on the CPU.
    Returns the running-mean film float32[H, W, 4] (updated in place if given)."""
    cam = np.ascontiguousarray(scene["camera"], np.float32)
    W, H = int(cam[19]), int(cam[20])
    if film is None:
        film = np.zeros((H, W, 4), np.float32)
    x0, y0, x1, y1 = window if window is not None else (0, 0, W, H)
    a = {k: np.ascontiguousarray(v) for k, v in scene.items()}
    lib().or_render(a["positions"].reshape(-1), a["normals"].reshape(-1), a["uvs"].reshape(-1),
                    a["triangles"].reshape(-1), a["tri_instance"], a["instances"].reshape(-1),
                    a["materials"].reshape(-1), a["lights"].reshape(-1), cam, a["render"],
                    len(a["positions"]), len(a["triangles"]), len(a["instances"]), len(a["materials"]),
                    len(a["lights"]), x0, y0, x1, y1, frame_start, spp, 1 if bvh else 0, threads, film)
    return film
