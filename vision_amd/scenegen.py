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
           up=(0, 1, 0), filter_param=None):
    fparam = {"radius": filter_radius, **(filter_param or {})}
    return {
        "type": kind,
        "param": {
            "fov_y": fov_y,
            "transform": {"type": "look_at", "param": {"position": list(position), "up": list(up),
                                                       "target_pos": list(target)}},
            "radiance_collector": {"type": "normal", "param": {"resolution": [width, height], "exposure": 1.0,
                                                               "accumulation": 1,
                                                               "tone_mapper": {"type": "linear"}}},
            "filter": {"type": filter_type, "param": fparam},
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


# ---------------------------------------------------------------- procedural scenes
def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: 20 * 4^level triangles (counter-clockwise seen from outside)."""
    t = (1 + 5 ** 0.5) / 2
    P = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mid = P[uniq[:, 0]] + P[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = len(P) + inv.reshape(3, -1)  # midpoint index of edges (01, 12, 20) per face
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        m01, m12, m20 = m[0], m[1], m[2]
        F = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1), np.stack([c, m20, m12], 1),
                            np.stack([m01, m12, m20], 1)])
        P = np.concatenate([P, mid])
    return P, F


def value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Seeded 3D lattice value noise in [-1, 1] with smoothstep interpolation."""
    i = np.floor(p).astype(np.int64)
    f = p - i
    w = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return (x & 0xFFFFFF).astype(np.float64) / float(0xFFFFFF) * 2 - 1

    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out = out + wx * wy * wz * h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    return out


def vertex_normals(P: np.ndarray, F: np.ndarray) -> np.ndarray:
    n = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    N = np.zeros_like(P)
    for k in range(3):
        np.add.at(N, F[:, k], n)
    return N / np.linalg.norm(N, axis=1, keepdims=True)


def write_vsmesh(path: str | Path, P: np.ndarray, F: np.ndarray, N: np.ndarray | None = None) -> Path:
    """Binary mesh: "VSMESH01", u32 nv, nt, flags (1: normals), f32 P, [f32 N], u32 tri."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as f:
        f.write(b"VSMESH01")
        f.write(np.array([len(P), len(F), 1 if N is not None else 0], "<u4").tobytes())
        f.write(np.asarray(P, "<f4").tobytes())
        if N is not None:
            f.write(np.asarray(N, "<f4").tobytes())
        f.write(np.asarray(F, "<u4").tobytes())
    return path


def sky_hdri(width: int = 2048, height: int = 1024, seed: int = 0, sun_peak: float = 5e4,
             sun_sigma_deg: float = 0.75) -> tuple[np.ndarray, tuple[float, float]]:
    """Seeded sky gradient plus one Gaussian sun lobe (peak radiance sun_peak)
    in the equirectangular layout of the spherical light: row v = theta / pi
    from the zenith (world +y), column u = phi / 2pi with phi measured from
    world +x toward +z.  Returns (float32 RGB[h, w, 3], (elevation, azimuth))."""
    rng = np.random.default_rng(seed)
    theta = (np.arange(height) + 0.5) / height * np.pi
    phi = (np.arange(width) + 0.5) / width * 2 * np.pi
    T, Ph = np.meshgrid(theta, phi, indexing="ij")
    up = np.cos(T)
    zenith = np.array([0.25, 0.45, 0.95]) * rng.uniform(0.9, 1.2)
    horizon = np.array([0.95, 0.95, 1.0]) * rng.uniform(1.5, 2.5)
    ground = np.array([0.12, 0.10, 0.08])
    k = np.clip(up, 0, 1)[..., None] ** 0.5
    sky = horizon * (1 - k) + zenith * k
    img = np.where(up[..., None] > 0, sky, ground * (1 + 0.3 * np.clip(-up, 0, 1))[..., None])
    elev = math.radians(rng.uniform(25, 50))
    azim = rng.uniform(0, 2 * np.pi)
    d = np.stack([np.sin(T) * np.cos(Ph), np.cos(T), np.sin(T) * np.sin(Ph)], -1)
    s = np.array([math.cos(elev) * math.cos(azim), math.sin(elev), math.cos(elev) * math.sin(azim)])
    ang = np.arccos(np.clip(d @ s, -1, 1))
    sig = math.radians(sun_sigma_deg)
    img = img + sun_peak * np.exp(-0.5 * (ang / sig) ** 2)[..., None] * np.array([1.0, 0.95, 0.85])
    return img.astype(np.float32), (elev, azim)


MATERIAL_FAMILIES = ("matte", "metal", "glass", "disney")


def procedural_scene(out_dir: str | Path, seed: int = 0, n_instances: int = 64, level: int = 5,
                     width: int = 1920, height: int = 1080, spp: int = 64, max_depth: int = 16,
                     spectrum: str = "hero", dimension: int = 4, hdri=(2048, 1024),
                     displacement: float = 0.1) -> dict:
    """Seeded instances of a displaced level-`level` icosphere (20 * 4^level
    triangles each, value-noise displacement of the given amplitude) on a
    2-triangle ground plane, materials drawn uniformly from matte, metal (GGX
    roughness U[0.05, 0.5], Cu / Au / Al), glass (Cauchy A = 1.5,
    B = 0.004 um^2) and Disney (principled), lit by a sky + sun HDRI.
    Meshes and the HDRI are written into out_dir; returns the scene dict
    (load it with base_dir = out_dir)."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    P0, F = icosphere(level)
    side = int(math.ceil(math.sqrt(n_instances)))
    spacing = 2.6
    shapes = [{"type": "quad", "name": "ground",
               "param": {"material": "Ground", "transform": xform(scale(400, 1, 400))}}]
    materials = [diffuse("Ground", (0.45, 0.45, 0.42))]
    for i in range(n_instances):
        mesh = out / f"sphere_{seed}_{level}_{i}.vsmesh"
        if not mesh.exists():
            noise = value_noise3(P0 * 2.5 + rng.uniform(-100, 100, 3), seed * 1000 + i)
            noise += 0.5 * value_noise3(P0 * 5.0 + rng.uniform(-100, 100, 3), seed * 1000 + i + 500)
            P = P0 * (1 + displacement * noise / 1.5)[:, None]
            write_vsmesh(mesh, P, F, vertex_normals(P, F))
        else:
            rng.uniform(-100, 100, 6)  # keep the stream aligned with a fresh generation
        gx, gz = i % side, i // side
        r = rng.uniform(0.7, 1.1)
        cx = (gx - (side - 1) / 2) * spacing + rng.uniform(-0.3, 0.3)
        cz = (gz - (side - 1) / 2) * spacing + rng.uniform(-0.3, 0.3)
        m = translate(cx, r * (1 + displacement), cz) @ rot_y(rng.uniform(0, 360)) @ scale(r, r, r)
        fam = MATERIAL_FAMILIES[int(rng.integers(0, 4))]
        name = f"M{i}"
        if fam == "matte":
            mat = {"type": "diffuse", "name": name, "param": {"color": [float(c) for c in rng.uniform(0.1, 0.9, 3)]}}
        elif fam == "metal":
            mat = {"type": "metal", "name": name,
                   "param": {"material_name": ["Cu", "Au", "Al"][int(rng.integers(0, 3))],
                             "roughness": float(rng.uniform(0.05, 0.5))}}
        elif fam == "glass":
            mat = {"type": "glass", "name": name,
                   "param": {"material_name": "Cauchy", "cauchy_a": 1.5, "cauchy_b": 0.004, "roughness": 0.05,
                             "color": [1.0, 1.0, 1.0]}}
        else:
            mat = {"type": "principled_bsdf", "name": name,
                   "param": {"color": [float(c) for c in rng.uniform(0.1, 0.9, 3)],
                             "metallic": float(rng.uniform(0, 1)), "roughness": float(rng.uniform(0.1, 0.8)),
                             "ior": 1.5, "spec_tint": [1.0, 1.0, 1.0],
                             "sheen_weight": float(rng.uniform(0, 0.5)),
                             "coat_weight": float(rng.uniform(0, 0.5)), "coat_roughness": 0.1,
                             "transmission_weight": 0.0}}
        materials.append(mat)
        shapes.append({"type": "model", "name": f"sphere{i}",
                       "param": {"fn": mesh.name, "material": name, "transform": xform(m)}})
    hdr = out / f"sky_{seed}_{hdri[0]}x{hdri[1]}.pfm"
    if not hdr.exists():
        img, _ = sky_hdri(hdri[0], hdri[1], seed)
        write_pfm(hdr, img)
    spec = {"type": "srgb"} if spectrum == "srgb" else {"type": "hero", "param": {"dimension": dimension}}
    return {
        "shapes": shapes,
        "materials": materials,
        "light_sampler": {"type": "power", "param": {"lights": [env_light(hdr.name)]}},
        "sampler": {"type": "independent", "param": {"spp": 1}},
        "spectrum": spec,
        "integrator": {"type": "pt", "param": {"min_depth": 5, "max_depth": max_depth, "rr_threshold": 1}},
        "camera": camera((0.0, 9.0, 23.0), (0.0, 0.6, 0.0), 38.0, width, height),
        "output": {"spp": spp, "fn": "procedural.exr"},
    }


SERVING_TRIANGLES = (100_000, 153_000, 234_000, 358_000, 547_000, 837_000, 1_280_000, 2_000_000)


def serving_scene(out_dir: str | Path, triangles: int, seed: int = 0, width: int = 256, height: int = 256,
                  max_depth: int = 8) -> dict:
    """The Cornell-box layout with about `triangles` triangles: the box, its
    two blocks and its light as in cornell_box, plus round(triangles / 20480)
    displaced level-5 icospheres (20 480 triangles each, value-noise
    displacement 0.1) of random diffuse albedo scattered through the box.
    Sphere meshes are shared by all scenes (written once into out_dir)."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    scene = cornell_box(width, height, max_depth=max_depth, light_sampler="power")
    n = max(1, int(round(triangles / 20480)))
    rng = np.random.default_rng(1000 + seed)
    P0, F = None, None
    for i in range(n):
        mesh = out / f"serve_sphere_{i}.vsmesh"
        if not mesh.exists():
            if P0 is None:
                P0, F = icosphere(5)
            r2 = np.random.default_rng(i)
            noise = value_noise3(P0 * 2.5 + r2.uniform(-100, 100, 3), 7000 + i)
            noise += 0.5 * value_noise3(P0 * 5.0 + r2.uniform(-100, 100, 3), 7500 + i)
            P = P0 * (1 + 0.1 * noise / 1.5)[:, None]
            write_vsmesh(mesh, P, F, vertex_normals(P, F))
        r = rng.uniform(0.06, 0.16)
        c = rng.uniform([-0.85, 0.15, -0.85], [0.85, 1.75, 0.85])
        m = translate(*c) @ rot_y(rng.uniform(0, 360)) @ scale(r, r, r)
        name = f"S{i}"
        scene["materials"].append(diffuse(name, tuple(float(x) for x in rng.uniform(0.1, 0.9, 3))))
        scene["shapes"].append({"type": "model", "name": f"sphere{i}",
                                "param": {"fn": mesh.name, "material": name, "transform": xform(m)}})
    return scene


def checker_rgba(width: int = 16, height: int = 8, a=(0.9, 0.2, 0.1), b=(0.1, 0.3, 0.8), cells: int = 4) -> np.ndarray:
    """A checkerboard image float32[H, W, 4] of two colours."""
    y, x = np.mgrid[0:height, 0:width]
    sel = ((x * cells // width) + (y * cells // height)) % 2 == 0
    img = np.where(sel[..., None], np.array(a, np.float32), np.array(b, np.float32))
    return np.concatenate([img, np.ones((height, width, 1), np.float32)], -1).astype(np.float32)


def textured_box(image_fn: str, width: int = 32, height: int = 32, color_space: str | None = None,
                 material: str = "diffuse", spectrum: str = "srgb", max_depth: int = 4) -> dict:
    """The Cornell box whose floor's colour slot is the image `image_fn`
    (an ImageNode, {"channels": "xyz", "node": {"type": "image", ...}})."""
    scene = cornell_box(width, height, max_depth=max_depth, light_sampler="power")
    param = {"fn": image_fn}
    if color_space:
        param["color_space"] = color_space
    node = {"channels": "xyz", "node": {"type": "image", "param": param}}
    if material == "diffuse":
        mat = {"type": "diffuse", "name": "Tex", "param": {"color": node}}
    else:
        mat = {"type": "principled_bsdf", "name": "Tex",
               "param": {"color": node, "roughness": 0.4, "metallic": 0.0, "spec_tint": node}}
    scene["materials"].append(mat)
    scene["shapes"][0]["param"]["material"] = "Tex"
    if spectrum != "srgb":
        scene["spectrum"] = {"type": "hero", "param": {"dimension": int(spectrum[-1])}}
    return scene


def write_scene(scene: dict, path: str | Path) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(scene, indent=1))
    return path


def restir_integrator(M_light: int = 10, M_bsdf: int = 1, spatial: bool = True, sample_num: int = 1,
                      radius: float = 30.0, temporal: bool = True, history_limit: int = 5, pairwise: bool = False,
                      temporal_mis: bool = False, extra_spatial_passes: int = 0, indirect: dict | None = None,
                      max_depth: int = 2, **extra) -> dict:
    """The "rt" integrator with its ReSTIR DI ("direct") block
    (src/render_core/integrators/rt.cpp, ReSTIR/direct.cpp:10-17,
    ReSTIR/common.h:19-67) and its ReSTIR GI ("indirect") block, closed
    unless `indirect` gives it (the reference opens it when the block is
    absent)."""
    direct = {"M_light": M_light, "M_bsdf": M_bsdf, "pairwise": pairwise,
              "spatial": {"open": spatial, "sample_num": sample_num, "radius": radius, "theta": 35, "depth": 0.3,
                          "extra_passes": extra_spatial_passes},
              "temporal": {"open": temporal, "history_limit": history_limit, "radius": 8, "theta": 35,
                           "depth": 0.3, "mis": temporal_mis}}
    direct.update(extra)
    gi = {"open": False} if indirect is None else indirect
    return {"type": "rt", "param": {"direct": direct, "indirect": gi, "max_depth": max_depth, "min_depth": 5,
                                    "rr_threshold": 1}}


def many_lights_box(width: int = 1920, height: int = 1080, n_quads: int = 5000, n_points: int = 256,
                    seed: int = 0, M_light: int = 32, M_bsdf: int = 1, sample_num: int = 5, radius: float = 30.0,
                    history_limit: int = 20, extra_spatial_passes: int = 1) -> dict:
    """The Cornell-box geometry lit by many lights for ReSTIR DI: `n_quads`
    small one-sided emissive quads (2 triangles each, radiance log-uniform in
    [1, 100], random colour and orientation) and `n_points` point and spot
    lights, chosen by the power light sampler; the ceiling light is removed.
    Defaults: 10 000 emissive triangles and 256 point/spot lights, 32 light
    candidates, temporal history 20, 5 spatial neighbours within 30 px in two
    passes."""
    rng = np.random.default_rng(seed)
    base = cornell_box(width, height)
    shapes = [s for s in base["shapes"] if s["name"] != "light"]
    for i in range(n_quads):
        pos = rng.uniform([-0.9, 0.15, -0.9], [0.9, 1.9, 0.9])
        size = rng.uniform(0.02, 0.06)
        radiance = float(np.exp(rng.uniform(np.log(1.0), np.log(100.0))))
        tint = rng.uniform(0.3, 1.0, 3)
        tint = tint / tint.max() * radiance
        m = translate(*pos) @ rot_y(rng.uniform(0, 360)) @ rot_x(rng.uniform(0, 360)) @ scale(size, 1, size)
        shapes.append({"type": "quad", "name": f"emitter{i}",
                       "param": {"material": "Light", "transform": xform(m),
                                 "emission": {"type": "area",
                                              "param": {"color": [float(c) for c in tint], "two_sided": False,
                                                        "scale": 1}}}})
    lights = []
    for i in range(n_points):
        pos = [float(v) for v in rng.uniform([-0.85, 0.2, -0.85], [0.85, 1.95, 0.85])]
        power = float(np.exp(rng.uniform(np.log(0.05), np.log(2.0))))
        color = [float(c) for c in rng.uniform(0.3, 1.0, 3)]
        if i % 2 == 0:
            lights.append({"type": "point", "param": {"position": pos, "color": color, "scale": power}})
        else:
            d = rng.normal(size=3)
            d[1] = -abs(d[1]) - 0.5
            lights.append({"type": "spot", "param": {"position": pos, "direction": [float(v) for v in d],
                                                     "angle": float(rng.uniform(20, 60)), "falloff": 10,
                                                     "color": color, "scale": power * 4}})
    base["shapes"] = shapes
    base["light_sampler"] = {"type": "power", "param": {"lights": lights}}
    base["integrator"] = restir_integrator(M_light=M_light, M_bsdf=M_bsdf, sample_num=sample_num, radius=radius,
                                           history_limit=history_limit,
                                           extra_spatial_passes=extra_spatial_passes)
    return base


def jitter_path(frames: int, position=(0, 1, 6.8), target=(0, 1, 0), amplitude: float = 0.01,
                seed: int = 0) -> list:
    """Camera-to-world matrices (4x4, column vectors) of a seeded jitter path:
    the look-at camera of the Cornell box with its position displaced by
    N(0, amplitude) per axis each frame."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(frames):
        p = np.asarray(position, np.float64) + rng.normal(scale=amplitude, size=3)
        f = np.asarray(target, np.float64) - p
        f /= np.linalg.norm(f)
        r = np.cross(f, [0.0, 1.0, 0.0])
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        m = np.eye(4)
        # the columns the scene loader's look_at builds: right, up, forward, position
        m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, f, p
        out.append(m)
    return out
