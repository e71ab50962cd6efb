"""Film output and the save chain.

The oracle's output buffer and final picture (radiance_collector.cpp:24-26,
normal.cpp:96-99, postprocessor.cpp:13-21, pipeline.cpp:313-330) are pinned
by closed forms of the tone mappers (impl.cpp:10-46) and the sRGB transfer
function (IEC 61966-2-1).  The library's image files are pinned by
independent readers and writers: PIL for PNG, numpy float16 for EXR halves,
and hand-built files that follow each format's specification (EXR ZIP
chunks, PNG filter types, Radiance run-length scanlines)."""
import math
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

import vision_amd as va
from oracle import oracle


# ---- tone mapping and sRGB -------------------------------------------------------

def test_output_buffer_closed_forms():
    ln2 = math.log(2.0)
    rt = np.array([[0, ln2, 2 * ln2, 1], [1e6, 0.5 * ln2, 0, 7]], np.float32)
    lin = oracle.film_output(rt, exposure=1.0, tone_mapper="linear")
    np.testing.assert_allclose(lin[0, :3], [0, 0.5, 0.75], atol=1e-7)
    np.testing.assert_allclose(lin[1, :3], [1, 1 - 2 ** -0.5, 0], atol=1e-7)
    assert np.all(lin[:, 3] == 1)  # w is forced to 1
    # exposure scales the radiance before the exponential
    np.testing.assert_allclose(oracle.film_output(rt, exposure=2.0)[0, 1], 0.75, atol=1e-7)
    # reinhard of 1 - exp(-v): 0.5 -> 1/3, 0.75 -> 3/7
    rh = oracle.film_output(rt, tone_mapper="reinhard")
    np.testing.assert_allclose(rh[0, 1:3], [1 / 3, 3 / 7], rtol=1e-6)
    # ACES fit (Narkowicz): x = 1 - exp(-v); at x = 0.5: (0.5(1.255+0.03)) / (0.5(1.215+0.59)+0.14)
    ac = oracle.film_output(rt, tone_mapper="aces")
    x = 0.5
    np.testing.assert_allclose(ac[0, 1], x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14), rtol=1e-6)
    assert ac[0, 0] == 0


def test_final_picture_closed_forms():
    out = np.array([[0.0, 0.18, 0.5, 0.3], [1.0, 0.0031308, 0.001, 0.0]], np.float32)
    pic = oracle.final_picture(out, "linear", gamma=True)
    # published sRGB values: 0.18 -> 0.46136, 0.5 -> 0.73536, 1 -> 1, linear segment 12.92 x
    np.testing.assert_allclose(pic[0, :3], [0, 0.4613561, 0.7353570], atol=2e-7)
    np.testing.assert_allclose(pic[1, [0, 2]], [1.0, 0.01292], atol=2e-7)
    # the two segments meet at the knee
    assert pic[1, 1] == pytest.approx(0.0404499, abs=1e-6)
    assert np.all(pic[:, 3] == 1)
    # decoding with the inverse transfer function gives the input back
    dec = np.where(pic[..., :3] <= 0.04045, pic[..., :3] / 12.92, ((pic[..., :3] + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(dec, out[..., :3], rtol=1e-5, atol=1e-7)
    # without gamma the linear tone mapper is the identity on RGB
    np.testing.assert_array_equal(oracle.final_picture(out, "linear", gamma=False)[..., :3], out[..., :3])
    # the tone mapper runs a second time on the output buffer: reinhard of 0.5 is 1/3
    np.testing.assert_allclose(oracle.final_picture(out, "reinhard", gamma=False)[0, 2], 1 / 3, rtol=1e-6)


# ---- image files ------------------------------------------------------------------

@pytest.fixture(scope="module")
def ctx():
    with va.Context(-1) as c:
        yield c


def _img(h=5, w=7, seed=0, lo=-2.0, hi=40.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (h, w, 4)).astype(np.float32)


def _parse_exr(data: bytes):
    """Minimal reader of a single-part scanline EXR, from the OpenEXR file
    layout document: magic, version, attributes, offset table, chunks."""
    magic, version = struct.unpack_from("<II", data, 0)
    assert magic == 20000630 and version & 0xFF == 2
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\0", name_end + 1)
        size = struct.unpack_from("<I", data, type_end + 1)[0]
        attrs[name] = data[type_end + 5:type_end + 5 + size]
        pos = type_end + 5 + size
    return attrs, pos + 1


def test_exr_round_trip_and_layout(ctx, tmp_path):
    a = _img()
    a[0, 0, 0] = np.inf
    a[1, 1, 1] = -0.0
    p = tmp_path / "x.exr"
    ctx.write_image(p, a)
    b = ctx.read_image(p)
    np.testing.assert_array_equal(b.view(np.uint32), a.view(np.uint32))
    data = p.read_bytes()
    attrs, table = _parse_exr(data)
    assert attrs["compression"] == b"\0"
    assert struct.unpack("<4i", attrs["dataWindow"]) == (0, 0, 6, 4)
    # channel list: A, B, G, R, all FLOAT (2), sampling 1
    chl = attrs["channels"]
    names = []
    q = 0
    while chl[q] != 0:
        e = chl.index(b"\0", q)
        names.append(chl[q:e].decode())
        ptype, _, xs, ys = struct.unpack_from("<iIii", chl, e + 1)
        assert (ptype, xs, ys) == (2, 1, 1)
        q = e + 17
    assert names == ["A", "B", "G", "R"]
    # one chunk per scanline: y, size, then the A, B, G, R rows
    off = struct.unpack_from("<5Q", data, table)
    for y in range(5):
        yy, size = struct.unpack_from("<iI", data, off[y])
        assert yy == y and size == 7 * 16
        row = np.frombuffer(data, np.float32, 28, off[y] + 8).reshape(4, 7)
        np.testing.assert_array_equal(row[::-1].T.view(np.uint32), a[y].view(np.uint32))


def _exr_zip_file(planes: dict, h: int, w: int, lines: int, compression: int) -> bytes:
    """An EXR with HALF channels compressed as the OpenEXR ZIP codec does:
    split even/odd bytes, byte-delta predictor, zlib."""
    names = sorted(planes)
    hdr = bytearray(struct.pack("<II", 20000630, 2))

    def attr(name, typ, value):
        hdr.extend(name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(value)) + value)

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iIii", 1, 0, 1, 1) for n in names) + b"\0"
    attr("channels", "chlist", chl)
    attr("compression", "compression", bytes([compression]))
    attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    attr("lineOrder", "lineOrder", b"\0")
    attr("pixelAspectRatio", "float", struct.pack("<f", 1))
    attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
    attr("screenWindowWidth", "float", struct.pack("<f", 1))
    hdr.append(0)
    chunks = []
    for y0 in range(0, h, lines):
        raw = b"".join(planes[n][y].astype("<f2").tobytes() for y in range(y0, min(h, y0 + lines)) for n in names)
        t = bytearray(len(raw))
        half = (len(raw) + 1) // 2
        t[:half], t[half:] = raw[0::2], raw[1::2]
        pred = bytearray(t)
        for i in range(1, len(t)):
            pred[i] = (t[i] - t[i - 1] + 128 + 256) & 0xFF
        z = zlib.compress(bytes(pred))
        chunks.append(struct.pack("<iI", y0, len(z)) + z)
    base = len(hdr) + 8 * len(chunks)
    offs, pos = [], base
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    return bytes(hdr) + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks)


@pytest.mark.parametrize("compression,lines", [(3, 16), (2, 1)], ids=["zip", "zips"])
def test_exr_reads_zip_half_channels(ctx, tmp_path, compression, lines):
    h, w = 37, 11
    rng = np.random.default_rng(4)
    planes = {c: rng.uniform(-3, 3000, (h, w)).astype(np.float16) for c in "RGB"}
    planes["R"][0, :3] = [np.float16(6e-8), np.float16(-0.0), np.float16(np.inf)]  # subnormal, signed zero, inf
    p = tmp_path / "z.exr"
    p.write_bytes(_exr_zip_file(planes, h, w, lines, compression))
    b = ctx.read_image(p)
    for i, c in enumerate("RGB"):
        np.testing.assert_array_equal(b[..., i], planes[c].astype(np.float32))
    assert np.all(b[..., 3] == 1)
    # a luminance-only file fills R, G and B
    y = {"Y": planes["G"]}
    p.write_bytes(_exr_zip_file(y, h, w, lines, compression))
    b = ctx.read_image(p)
    for i in range(3):
        np.testing.assert_array_equal(b[..., i], planes["G"].astype(np.float32))


def test_hdr_round_trip_and_rle(ctx, tmp_path):
    a = _img(lo=0.0, hi=50.0, w=9)
    a[2, 3, :3] = 0
    p = tmp_path / "x.hdr"
    ctx.write_image(p, a)
    b = ctx.read_image(p)
    # RGBE keeps 8 mantissa bits of the largest channel
    m = a[..., :3].max(axis=-1, keepdims=True)
    assert np.all(np.abs(b[..., :3] - a[..., :3]) <= m / 128 + 1e-30)
    assert np.all(b[2, 3, :3] == 0)
    # a new-style run-length file (Ward's scanline encoding) decodes to the same pixels
    rgbe = np.frombuffer(p.read_bytes()[-9 * 5 * 4:], np.uint8).reshape(5, 9, 4)
    body = bytearray()
    for y in range(5):
        body += bytes([2, 2, 0, 9])
        for c in range(4):
            ch = rgbe[y, :, c]
            body += bytes([128 + 3, ch[0]])          # a run of three copies of the first byte
            body += bytes([6]) + ch[3:9].tobytes()    # then six literals
    rle = rgbe.copy()
    rle[:, 1:3] = rle[:, :1]
    q = tmp_path / "r.hdr"
    q.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 5 +X 9\n" + bytes(body))
    c = ctx.read_image(q)
    e = rle[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    np.testing.assert_array_equal(c[..., :3], (rle[..., :3] * f[..., None]).astype(np.float32))


def test_png_write_matches_pil(ctx, tmp_path):
    a = _img(lo=-0.2, hi=1.2, h=6, w=13)
    a[0, 0] = np.nan
    p = tmp_path / "x.png"
    ctx.write_image(p, a)
    with Image.open(p) as im:
        assert im.mode == "RGBA" and im.size == (13, 6)
        got = np.asarray(im)
    want = np.where(np.isnan(a), 0, np.floor(np.clip(np.nan_to_num(a), 0, 1) * 255 + 0.5)).astype(np.uint8)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(ctx.read_image(p), want.astype(np.float32) / 255)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16"])
def test_png_read_matches_pil(ctx, tmp_path, mode):
    rng = np.random.default_rng(9)
    rgb = rng.integers(0, 256, (17, 23, 3), dtype=np.uint8)
    base = Image.fromarray(rgb, "RGB")
    if mode == "I;16":
        im = Image.fromarray(rng.integers(0, 65536, (17, 23), dtype=np.uint16))
    elif mode == "P":
        im = base.convert("P", palette=Image.Palette.ADAPTIVE, colors=64)
    else:
        im = base.convert(mode)
    p = tmp_path / "m.png"
    im.save(p)
    got = ctx.read_image(p)
    with Image.open(p) as r:
        if mode == "I;16":
            want = np.asarray(r).astype(np.float32) / 65535
            want = np.stack([want] * 3 + [np.ones_like(want)], -1)
        else:
            want = np.asarray(r.convert("RGBA")).astype(np.float32) / 255
    np.testing.assert_allclose(got, want, atol=1e-7)


def _png_chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))


def test_png_filters(ctx, tmp_path):
    """Rows filtered with each of the five PNG filter types (ISO/IEC 15948
    sec. 9) decode to the original bytes."""
    rng = np.random.default_rng(2)
    h, w, bpp = 10, 9, 3
    px = rng.integers(0, 256, (h, w * bpp)).astype(np.int32)
    raw = bytearray()
    for y in range(h):
        f = y % 5
        row = px[y]
        up = px[y - 1] if y else np.zeros_like(row)
        left = np.concatenate([np.zeros(bpp, np.int32), row[:-bpp]])
        ul = np.concatenate([np.zeros(bpp, np.int32), up[:-bpp]])
        if f == 0:
            pred = np.zeros_like(row)
        elif f == 1:
            pred = left
        elif f == 2:
            pred = up
        elif f == 3:
            pred = (left + up) // 2
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
        raw += bytes([f]) + ((row - pred) & 0xFF).astype(np.uint8).tobytes()
    data = (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
            _png_chunk(b"IDAT", zlib.compress(bytes(raw))) + _png_chunk(b"IEND", b""))
    p = tmp_path / "f.png"
    p.write_bytes(data)
    got = ctx.read_image(p)
    np.testing.assert_array_equal(got[..., :3], px.reshape(h, w, 3).astype(np.float32) / 255)
    # a corrupted CRC is an error, not a silent read
    bad = bytearray(data)
    bad[30] ^= 1
    p.write_bytes(bytes(bad))
    with pytest.raises(va.VisionError):
        ctx.read_image(p)


def test_pfm_round_trip(ctx, tmp_path):
    a = _img()
    a[..., 3] = 1
    p = tmp_path / "x.pfm"
    ctx.write_image(p, a)
    np.testing.assert_array_equal(ctx.read_image(p), a)
    # rows are stored bottom to top
    raw = np.frombuffer(p.read_bytes()[-5 * 7 * 12:], "<f4").reshape(5, 7, 3)
    np.testing.assert_array_equal(raw[::-1], a[..., :3])


def test_image_errors(ctx, tmp_path):
    a = np.zeros((2, 2, 4), np.float32)
    with pytest.raises(va.VisionError) as e:
        ctx.write_image(tmp_path / "x.jpg", a)
    assert e.value.status == 3  # unsupported
    with pytest.raises(va.VisionError):
        ctx.read_image(tmp_path / "missing.exr")
    p = tmp_path / "t.exr"
    ctx.write_image(p, np.ones((4, 4, 4), np.float32))
    p.write_bytes(p.read_bytes()[:-20])
    with pytest.raises(va.VisionError):
        ctx.read_image(p)
