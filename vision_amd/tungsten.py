"""Tungsten scene -> Vision scene JSON (tools/tungsten2vision.py of the
reference, the converter that produced its cbox / staircase / spaceship
scenes).

    python -m vision_amd.tungsten tungsten_scene.json [vision_scene.json]

The mapping follows the reference tool (tools/tungsten2vision.py:53-451):

* bsdfs (:266-284): lambert / oren_nayar -> a diffuse material with the
  albedo; dielectric / rough_dielectric -> glass {ior, roughness, color};
  mirror -> mirror {color, roughness 0.001}; conductor / rough_conductor ->
  metal {material_name, eta, k, roughness, no remapping}; plastic /
  rough_plastic -> substrate {color, ior 1.5, specular 0.04, roughness};
  null -> black diffuse; anything else -> the reference's fixed Disney-like
  stack.  The reference tool names the matte and Disney materials "matte"
  and "disney", which its own loader no longer registers (its plugins are
  diffuse.cpp and principled_bsdf.cpp); this converter writes "diffuse" and
  "principled_bsdf" with the same parameters, so the output loads.
* primitives (:286-315): quad and disk -> quad, cube -> cube, mesh -> model
  (the .wo3 name with .obj, the mesh pre-converted), each with a matrix4x4
  transform T * Ry * Rx * Rz * S from the Tungsten transform (rotation in
  degrees applied z, then x, then y: rotateZXY, :213-239), the matrix
  written column by column as the reference's loader reads it; inline
  bsdfs become named materials; emission -> an area light of that radiance,
  power P -> radiance P / (100 pi) at scale 30 (:241-252); infinite_sphere /
  cap / skydome -> a white environment light in the light sampler.
* camera (:163-197): a thin lens (radius 0, i.e. a pinhole) at look_at of
  the Tungsten transform, its fov, resolution and a box filter of radius
  0.5; integrator (:199-209): pt with Tungsten's min / max bounces; sampler
  independent; light sampler uniform over the environment lights.
"""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path

import numpy as np


def _vec(value, dim=3):
    """convert_vec (:41-50): scalars broadcast, strings become image slots."""
    if isinstance(value, str):
        return {"fn": value, "color_space": "srgb"}
    if not isinstance(value, list):
        return [value] * dim
    if len(value) != dim:
        raise ValueError(f"expected {dim} components, got {value}")
    return value


def _rot(axis: int, deg: float) -> np.ndarray:
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def shape_matrix(transform: dict) -> np.ndarray:
    """T * rotateZXY(R) * S (convert_srt, :213-217), a 4x4 acting on column
    vectors."""
    t = np.eye(4)
    t[:3, 3] = transform.get("position", [0, 0, 0])
    r = transform.get("rotation", [0, 0, 0])
    s = np.diag(list(map(float, transform.get("scale", [1, 1, 1]))) + [1.0])
    return t @ _rot(1, r[1]) @ _rot(0, r[0]) @ _rot(2, r[2]) @ s


def _transform(transform: dict) -> dict:
    m = shape_matrix(transform)
    return {"type": "matrix4x4", "param": {"matrix4x4": [[float(v) for v in m[:, c]] for c in range(4)]}}


def convert_material(m: dict) -> dict | None:
    t = m["type"]
    name = m["name"]
    if t in ("lambert", "oren_nayar"):
        return {"type": "diffuse", "name": name, "param": {"color": _vec(m["albedo"], 3)}}
    if t in ("dielectric", "rough_dielectric"):
        return {"type": "glass", "name": name,
                "param": {"ior": m["ior"], "roughness": _vec(m.get("roughness", 0), 2),
                          "color": _vec(m.get("albedo", 1), 3)}}
    if t == "mirror":
        return {"type": "mirror", "name": name, "param": {"color": _vec(m.get("albedo", 1), 3), "roughness": 0.001}}
    if t in ("conductor", "rough_conductor"):
        return {"type": "metal", "name": name,
                "param": {"material_name": m.get("material", ""), "eta": _vec(m.get("eta", 0), 3),
                          "k": _vec(m.get("k", 0), 3), "roughness": _vec(m.get("roughness", 0.01), 2),
                          "remapping_roughness": False}}
    if t in ("plastic", "rough_plastic"):
        return {"type": "substrate", "name": name,
                "param": {"color": _vec(m.get("albedo", 1), 3), "ior": 1.5, "specular": [0.04, 0.04, 0.04],
                          "roughness": _vec(m.get("roughness", 0.01), 2), "remapping_roughness": False}}
    if t == "null":
        return {"type": "diffuse", "name": name, "param": {"color": [0, 0, 0]}}
    # every other Tungsten bsdf: the reference's fixed Disney-like stack
    # (convert_disney, :140-161) in principled_bsdf's parameter names
    return {"type": "principled_bsdf", "name": name,
            "param": {"color": _vec(m.get("albedo", 1), 3), "ior": 1.5, "roughness": 0.1, "metallic": 0.0,
                      "spec_tint": [0.9, 0.9, 0.9], "anisotropic": 0.0, "sheen_weight": 0.0,
                      "coat_weight": 0.59, "coat_roughness": 0.38, "transmission_weight": 0.0}}


def _emission(shape: dict) -> tuple[list, float]:
    """get_emission (:241-248)."""
    if "emission" in shape:
        return _vec(shape["emission"], 3), 1.0
    p = np.asarray(_vec(shape["power"], 3), np.float64) / (100.0 * math.pi)
    return [float(v) for v in p], 30.0


def convert(scene: dict) -> dict:
    """Convert a parsed Tungsten scene to a Vision scene dict."""
    lights: list[dict] = []
    materials = [m for m in (convert_material(b) for b in scene.get("bsdfs", [])) if m]
    shapes = []
    for i, prim in enumerate(scene.get("primitives", [])):
        t = prim["type"]
        out = None
        if t in ("quad", "disk"):
            out = {"type": "quad", "name": f"shape_{i}",
                   "param": {"width": 1.0, "height": 1.0, "material": prim.get("bsdf"),
                             "transform": _transform(prim.get("transform", {}))}}
        elif t == "cube":
            out = {"type": "cube", "name": f"shape_{i}",
                   "param": {"x": 1, "y": 1, "z": 1, "material": prim.get("bsdf"),
                             "transform": _transform(prim.get("transform", {}))}}
        elif t == "mesh":
            bsdf = prim.get("bsdf")
            out = {"type": "model", "name": f"shape_{i}",
                   "param": {"fn": prim["file"][:-4] + ".obj", "smooth": prim.get("smooth", True),
                             "material": bsdf,
                             "transform": _transform(prim.get("transform", {}))}}
        if t in ("infinite_sphere", "infinite_sphere_cap", "skydome"):
            # convert_envmap (:254-268): a white environment light
            lights.append({"type": "environment",
                           "param": {"texture": [1, 1, 1], "o2w": {"type": "Euler", "param": {"yaw": 0}},
                                     "scale": 1}})
        elif out is not None and ("emission" in prim or "power" in prim):
            color, scale = _emission(prim)
            out["param"]["emission"] = {"type": "area", "param": {"color": color, "two_sided": False, "scale": scale}}
        if out is None:
            continue
        if isinstance(out["param"]["material"], dict):
            # an inline bsdf becomes a named material
            mat = dict(out["param"]["material"])
            mat["name"] = out["name"] + "_material"
            out["param"]["material"] = mat["name"]
            conv = convert_material(mat)
            if conv:
                materials.append(conv)
        shapes.append(out)
    cam = scene["camera"]
    tr = cam["transform"]
    integ = scene.get("integrator", {})
    return {
        "shapes": shapes,
        "materials": materials,
        "sampler": {"type": "independent", "name": "sampler0", "param": {"spp": 1}},
        "camera": {"type": "thin_lens",
                   "param": {"fov_y": cam["fov"], "name": "view1", "velocity": 5,
                             "transform": {"type": "look_at",
                                           "param": {"position": tr["position"], "up": tr["up"],
                                                     "target_pos": tr["look_at"]}},
                             "film": {"type": "rgb", "param": {"resolution": _vec(cam.get("resolution", [768, 768]), 2),
                                                               "fb_state": 0}},
                             "filter": {"type": "box", "name": "boxFilter", "param": {"radius": [0.5, 0.5]}}}},
        "warper": {"type": "alias_table"},
        "integrator": {"type": "pt", "param": {"min_depth": integ.get("min_bounces", 0),
                                               "max_depth": integ.get("max_bounces", 64), "rr_threshold": 1}},
        "light_sampler": {"type": "uniform", "param": {"lights": lights}},
        "output": {"spp": 1024, "fn": "output.png"},
    }


def convert_file(src: str | Path, dst: str | Path | None = None) -> Path:
    src = Path(src)
    dst = Path(dst) if dst is not None else src.with_name("vision_scene.json")
    dst.write_text(json.dumps(convert(json.loads(src.read_text())), indent=4))
    return dst


def main(argv: list[str]) -> int:
    if len(argv) < 2:
        print(__doc__.splitlines()[3].strip())
        return 2
    out = convert_file(argv[1], argv[2] if len(argv) > 2 else None)
    print(f"vision scene saved to {out.resolve()}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
