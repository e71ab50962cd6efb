"""Convert the reference's numeric data tables to little-endian float32 files.

Sources (data, not code) in the reference tree:
  src/base/scattering/precomputed_table.h   energy-compensation LUTs
  src/render_core/materials/ltc_sheen_table.h  sheen LTC fits (Zeltner 2022)
  src/base/color/cie.h                      CIE 1931 X, Y, Z and D65 at 1 nm, 360-830 nm
  src/render_core/materials/metal_ior.inl.h measured metal eta/k (95 samples, 360-830 nm)
Run once where the reference is mounted:  python vision_amd/data/extract.py /root/reference
"""
import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def block(text, start_pat):
    i = text.index(start_pat)
    j = text.index("};", i)
    body = text[text.index("{", i) + 1:j]
    body = re.sub(r"float[234]?\(", "(", body)
    body = body.replace("make_float2(", "(")
    vals = [float(v.rstrip("f")) for v in re.findall(NUM + "f?", body)]
    return np.array(vals, np.float32)


def main(ref):
    ref = Path(ref)
    t = (ref / "src/base/scattering/precomputed_table.h").read_text()
    out = {
        "lut_dielectric": (block(t, "DielectricLobe_Table["), 32768 * 2),
        "lut_dielectric_inv": (block(t, "DielectricInvLobe_Table["), 32768 * 2),
        "lut_specular": (block(t, "SpecularLobe_Table["), 32768),
        "lut_coat": (block(t, "CoatLobe_Table["), 32768),
        "lut_pure_reflection": (block(t, "PureReflectionLobe_Table["), 1024),
    }
    s = (ref / "src/render_core/materials/ltc_sheen_table.h").read_text()
    out["lut_sheen_volume"] = (block(s, "SheenLTCTableVolume["), 4096)
    out["lut_sheen_approx"] = (block(s, "SheenLTCTableApprox["), 4096)
    c = (ref / "src/base/color/cie.h").read_text()
    cie = np.concatenate([block(c, "> X ="), block(c, "> Y ="), block(c, "> Z ="), block(c, "> D65 =")])
    out["cie_xyz_d65"] = (cie, 471 * 4)
    m = (ref / "src/render_core/materials/metal_ior.inl.h").read_text()
    names = re.findall(r"ocarina::array (\w+)\{", m)
    for n in names:
        v = block(m, f"ocarina::array {n}{{")
        out[f"metal_{n}"] = (v, len(v))
    for name, (arr, n) in out.items():
        assert arr.size == n, (name, arr.size, n)
        arr.astype("<f4").tofile(OUT / f"{name}.f32")
        print(name, arr.size)
    (OUT / "metals.txt").write_text("\n".join(names) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
