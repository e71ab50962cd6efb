"""Synthetic Vision-schema scenes for the benchmark and tests.

Scenes are plain dicts in the reference's scene-JSON schema (shapes,
materials, light_sampler, camera, integrator, ... ; src/core/scene_desc.cpp:37-53),
so they load through the same path as the reference's own scene files.
Matrices are written as "matrix4x4" transforms listing four columns.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np


def _cols(m: np.ndarray) -> list:
    """4x4 matrix -> the scene format's list of columns."""
    return [[float(m[r, c]) for r in range(4)] for c in range(4)]


def translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scale(x, y, z):
    return np.diag([x, y, z, 1.0])


def rot_x(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(4)
    m[1, 1], m[1, 2], m[2, 1], m[2, 2] = c, -s, s, c
    return m


def rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(4)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    return m


def rot_z(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(4)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, -s, s, c
    return m


def xform(m: np.ndarray) -> dict:
    return {"type": "matrix4x4", "param": {"matrix4x4": _cols(m)}}


def diffuse(name, color):
    return {"type": "diffuse", "name": name, "param": {"color": [float(c) for c in color]}}


def camera(position, target, fov_y, width, height, kind="pinhole", filter_type="box", filter_radius=0.5,
           up=(0, 1, 0)):
    return {
        "type": kind,
        "param": {
            "fov_y": fov_y,
            "transform": {"type": "look_at", "param": {"position": list(position), "up": list(up),
                                                       "target_pos": list(target)}},
            "radiance_collector": {"type": "normal", "param": {"resolution": [width, height], "exposure": 1.0,
                                                               "accumulation": 1,
                                                               "tone_mapper": {"type": "linear"}}},
            "filter": {"type": filter_type, "param": {"radius": filter_radius}},
        },
    }


def cornell_box(width=256, height=256, spp=16, max_depth=8, min_depth=5, light_sampler="uniform") -> dict:
    """Cornell-box layout: 5 diffuse wall quads (0.73 gray; 0.65/0.05/0.05
    red; 0.12/0.45/0.15 green), two boxes rotated +-17 degrees (24 triangles)
    and one quad area light of radiance (17, 12, 4); pinhole, box filter, sRGB."""
    gray, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    shapes = [
        {"type": "quad", "name": "floor", "param": {"material": "Gray", "transform": xform(scale(2, 1, 2))}},
        {"type": "quad", "name": "ceiling",
         "param": {"material": "Gray", "transform": xform(translate(0, 2, 0) @ rot_x(180) @ scale(2, 1, 2))}},
        {"type": "quad", "name": "back",
         "param": {"material": "Gray", "transform": xform(translate(0, 1, -1) @ rot_x(90) @ scale(2, 1, 2))}},
        {"type": "quad", "name": "left",
         "param": {"material": "Red", "transform": xform(translate(-1, 1, 0) @ rot_z(-90) @ scale(2, 1, 2))}},
        {"type": "quad", "name": "right",
         "param": {"material": "Green", "transform": xform(translate(1, 1, 0) @ rot_z(90) @ scale(2, 1, 2))}},
        {"type": "cube", "name": "short_box",
         "param": {"x": 1, "y": 1, "z": 1, "material": "Gray",
                   "transform": xform(translate(0.33, 0.3, 0.37) @ rot_y(-17) @ scale(0.6, 0.6, 0.6))}},
        {"type": "cube", "name": "tall_box",
         "param": {"x": 1, "y": 1, "z": 1, "material": "Gray",
                   "transform": xform(translate(-0.335, 0.6, -0.29) @ rot_y(17) @ scale(0.6, 1.2, 0.6))}},
        {"type": "quad", "name": "light",
         "param": {"width": 1, "height": 1, "material": "Light",
                   "transform": xform(translate(0, 1.98, 0) @ scale(0.47, -1, -0.38)),
                   "emission": {"type": "area", "param": {"color": {"channels": "xyz", "node": [17, 12, 4]},
                                                          "two_sided": False, "scale": 1}}}},
    ]
    return {
        "shapes": shapes,
        "materials": [diffuse("Gray", gray), diffuse("Red", red), diffuse("Green", green),
                      diffuse("Light", (0, 0, 0))],
        "sampler": {"type": "independent", "param": {"spp": 1}},
        "light_sampler": {"type": light_sampler, "param": {}},
        "spectrum": {"type": "srgb"},
        "integrator": {"type": "pt", "param": {"min_depth": min_depth, "max_depth": max_depth, "rr_threshold": 1}},
        "camera": camera((0, 1, 6.8), (0, 1, 0), 19.5, width, height),
        "output": {"spp": spp, "fn": "cbox.exr"},
    }


def furnace_box(albedo=(0.5, 0.5, 0.5), radiance=(1.0, 1.0, 1.0), width=32, height=32, max_depth=64,
                material_sigma=None) -> dict:
    """Closed diffuse box emitting `radiance` on every inner wall, camera at
    its centre: every pixel converges to radiance / (1 - albedo)."""
    mat = diffuse("Wall", albedo)
    if material_sigma is not None:
        mat["param"]["sigma"] = material_sigma
    return {
        "shapes": [{"type": "cube", "name": "box",
                    "param": {"x": 2, "y": 2, "z": 2, "material": "Wall", "transform": xform(np.eye(4)),
                              "emission": {"type": "area",
                                           "param": {"color": list(radiance), "two_sided": True, "scale": 1}}}}],
        "materials": [mat],
        "light_sampler": {"type": "uniform", "param": {}},
        "integrator": {"type": "pt", "param": {"min_depth": 5, "max_depth": max_depth, "rr_threshold": 1}},
        "camera": camera((0, 0, 0.3), (0, 0, -1), 60, width, height),
    }


def lit_plane(light: dict, albedo=0.5, cam_height=1.0, fov=2.0, width=16, height=16, extra_shapes=()) -> dict:
    """Large diffuse floor (y = 0) seen straight down from (0, cam_height, 0)
    through a narrow camera, lit by `light` (a light_sampler light or an
    emissive shape); max_depth 2, so only direct lighting reaches the film."""
    shapes = [{"type": "quad", "name": "floor",
               "param": {"material": "Floor", "transform": xform(scale(200, 1, 200))}}]
    shapes += list(extra_shapes)
    lights = [light] if light else []
    return {
        "shapes": shapes,
        "materials": [diffuse("Floor", (albedo, albedo, albedo)), diffuse("Black", (0, 0, 0))],
        "light_sampler": {"type": "uniform", "param": {"lights": lights}},
        "integrator": {"type": "pt", "param": {"min_depth": 5, "max_depth": 2, "rr_threshold": 1}},
        "camera": camera((0, cam_height, 0), (0, 0, 0), fov, width, height, up=(0, 0, -1)),
    }


def write_pfm(path: str | Path, rgb: np.ndarray) -> Path:
    """Write float32 RGB[h, w, 3] (top row first) as a little-endian PFM,
    whose rows run bottom to top."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    img = np.ascontiguousarray(np.asarray(rgb, np.float32)[::-1])
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(img.astype("<f4").tobytes())
    return path


def env_light(fn: str, scale: float = 1.0, o2w: dict | None = None, flip_u: bool = False) -> dict:
    """A spherical (HDRI) environment light over image file `fn`."""
    p = {"color": {"channels": "xyz", "node": {"fn": fn, "color_space": "linear"}}, "scale": scale,
         "flip_u": flip_u}
    if o2w is not None:
        p["o2w"] = o2w
    return {"type": "spherical", "param": p}


def write_scene(scene: dict, path: str | Path) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(scene, indent=1))
    return path
