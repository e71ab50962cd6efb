"""Write the Tungsten fixtures from the reference's Cornell-box files
(res/render_scene/cbox/tungsten_scene.json and TungstenRender.png, an image
of that scene rendered by the Tungsten renderer, 8-bit with its "gamma" tone
map): the scene file as it is, and the render's 32 x 32 block means of the
8-bit codes (float32 [32, 32, 3]; the 1024^2 image is too large to keep).  Run in the build container, where the reference is
mounted:

    python tests/golden/make_tungsten_fixtures.py /root/reference
"""
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))


def main(ref_root: str) -> None:
    from PIL import Image
    src = Path(ref_root) / "res" / "render_scene" / "cbox"
    shutil.copyfile(src / "tungsten_scene.json", HERE / "tungsten_cbox.json")
    img = np.asarray(Image.open(src / "TungstenRender.png").convert("RGB")).astype(np.float64)
    h, w = img.shape[:2]
    blocks = img.reshape(32, h // 32, 32, w // 32, 3).mean(axis=(1, 3)).astype(np.float32)
    np.save(HERE / "tungsten_cbox_render_32x32.npy", blocks, allow_pickle=False)
    print(f"{w}x{h} render, mean {img.mean(axis=(0, 1))}, blocks {blocks.shape}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
