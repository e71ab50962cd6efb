"""The JPEG reader (vision_amd/csrc/jpeg.cpp) against an independent decoder,
PIL's libjpeg: baseline and progressive files, 4:4:4 / 4:2:2 / 4:2:0
chroma, grey, odd sizes, restart intervals, high and low quality.  Textures
the reference's scenes name are mostly .jpg (ImagePool::load_texture,
src/base/mgr/image_pool.cpp:13-34).

Tolerance: this reader's IDCT runs in double precision, libjpeg's default
one in 13-bit fixed point, so a sample can round one code apart and the
colour conversion can carry that to +-3 codes in R or B; the mean absolute
difference stays below 0.25 codes per channel."""
import io
import json
from pathlib import Path

import numpy as np
import pytest
from PIL import Image

import vision_amd as va
from vision_amd import scenegen


@pytest.fixture(scope="module")
def ctx():
    with va.Context(-1) as c:
        yield c


def _picture(w, h, seed=0):
    """A smooth gradient with sharp edges and noise, so every coefficient
    band and both chroma planes carry signal."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 127 + 120 * np.sin(x / 7.0) * np.cos(y / 11.0)
    g = 255 * (x / max(w - 1, 1))
    b = np.where((x // 8 + y // 8) % 2 == 0, 230.0, 20.0)
    img = np.stack([r, g, b], -1) + rng.normal(0, 12, (h, w, 3))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _compare(ctx, tmp_path, data: bytes, name="t.jpg"):
    p = tmp_path / name
    p.write_bytes(data)
    got = ctx.read_image(p)
    with Image.open(io.BytesIO(data)) as im:
        ref = np.asarray(im.convert("RGB")).astype(np.float64)
    assert got.shape[:2] == ref.shape[:2]
    np.testing.assert_array_equal(got[..., 3], 1.0)
    diff = np.abs(np.rint(got[..., :3].astype(np.float64) * 255) - ref)
    return diff.max(), diff.mean()


@pytest.mark.parametrize("progressive", [False, True])
@pytest.mark.parametrize("subsampling", [0, 1, 2])  # 4:4:4, 4:2:2, 4:2:0
@pytest.mark.parametrize("quality", [95, 40])
def test_colour_files_match_libjpeg(ctx, tmp_path, progressive, subsampling, quality):
    buf = io.BytesIO()
    Image.fromarray(_picture(203, 117, seed=subsampling)).save(
        buf, "JPEG", quality=quality, subsampling=subsampling, progressive=progressive)
    mx, mean = _compare(ctx, tmp_path, buf.getvalue())
    print(f"progressive={progressive} subsampling={subsampling} q={quality}: max {mx}, mean {mean:.4f}")
    assert mx <= 3 and mean < 0.25


@pytest.mark.parametrize("progressive", [False, True])
def test_grey_files_match_libjpeg(ctx, tmp_path, progressive):
    grey = _picture(64, 33)[..., 1]
    buf = io.BytesIO()
    Image.fromarray(grey, "L").save(buf, "JPEG", quality=85, progressive=progressive)
    mx, mean = _compare(ctx, tmp_path, buf.getvalue())
    assert mx <= 1 and mean < 0.1


def test_restart_intervals_and_optimised_tables(ctx, tmp_path):
    buf = io.BytesIO()
    Image.fromarray(_picture(150, 90, seed=7)).save(buf, "JPEG", quality=90, optimize=True,
                                                    restart_marker_blocks=3)
    data = buf.getvalue()
    assert b"\xff\xdd" in data and b"\xff\xd0" in data  # DRI segment and RST markers present
    mx, mean = _compare(ctx, tmp_path, data)
    assert mx <= 3 and mean < 0.25


def test_tiny_image_and_single_chroma_column(ctx, tmp_path):
    for w, h in [(1, 1), (2, 1), (3, 5), (17, 9)]:
        buf = io.BytesIO()
        Image.fromarray(_picture(w, h, seed=w)).save(buf, "JPEG", quality=90, subsampling=2)
        mx, _ = _compare(ctx, tmp_path, buf.getvalue())
        assert mx <= 3, (w, h)


def test_corrupt_and_unsupported_files_fail_loudly(ctx, tmp_path):
    buf = io.BytesIO()
    Image.fromarray(_picture(32, 32)).save(buf, "JPEG", quality=90)
    data = buf.getvalue()
    (tmp_path / "cut.jpg").write_bytes(data[:200])
    with pytest.raises(va.VisionError):
        ctx.read_image(tmp_path / "cut.jpg")
    cmyk = io.BytesIO()
    Image.fromarray(_picture(16, 16)).convert("CMYK").save(cmyk, "JPEG")
    (tmp_path / "cmyk.jpg").write_bytes(cmyk.getvalue())
    with pytest.raises(va.VisionError, match="component count"):
        ctx.read_image(tmp_path / "cmyk.jpg")


REF_RES = Path("/root/reference/res")


@pytest.mark.skipif(not REF_RES.is_dir(), reason="the reference's scene resources are not on this machine")
def test_reference_scene_textures_match_libjpeg(ctx, tmp_path):
    """Every distinct .jpg under the reference's res/ (baseline and
    progressive files from other encoders), read in place."""
    seen, n = set(), 0
    for p in sorted(REF_RES.rglob("*.jpg")):
        key = (p.name, p.stat().st_size)
        if key in seen or p.stat().st_size > 2_000_000:
            continue
        seen.add(key)
        data = p.read_bytes()
        try:
            with Image.open(io.BytesIO(data)) as im:
                im.load()
        except OSError:
            continue  # a file libjpeg cannot read either (truncated)
        mx, mean = _compare(ctx, tmp_path, data, p.name)
        print(f"{p.relative_to(REF_RES)}: max {mx}, mean {mean:.4f}")
        assert mx <= 4 and mean < 0.3, p
        n += 1
    assert n >= 10


def test_jpeg_texture_is_srgb_decoded(ctx, tmp_path):
    buf = io.BytesIO()
    Image.fromarray(_picture(24, 12, seed=3)).save(buf, "JPEG", quality=92)
    (tmp_path / "paint.jpg").write_bytes(buf.getvalue())
    ex = ctx.load_text(json.dumps(scenegen.textured_box("paint.jpg")), str(tmp_path)).build_accel().export()
    np.testing.assert_array_equal(ex["texture_info"], [[0, 24, 12, 0]])
    codes = ctx.read_image(tmp_path / "paint.jpg")[..., :3].astype(np.float64)
    lin = np.where(codes <= 0.04045, codes / 12.92, ((codes + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(ex["textures"].reshape(12, 24, 4)[..., :3], lin, rtol=2e-6, atol=1e-7)
