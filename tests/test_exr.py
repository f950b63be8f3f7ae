"""OpenEXR decoding of environment maps (hpt_exr_info / hpt_exr_decode_rgb,
host code, no device).  The reference's configurations name OpenEXR HDRIs
(conf/dragon.yaml:13 and siblings); no OpenEXR library or file exists in this
image, so the decoder is pinned by files assembled here byte by byte from the
format, by every half-float bit pattern against numpy's conversion, and by
round trips through the independent writer in exr_encode.py over each
compression, pixel type, odd sizes, both PIZ wavelet widths and raw-stored
chunks.  Malformed and unsupported files fail with a message, never a crash."""
import os
import struct
import zlib

import numpy as np
import pytest

import exr_encode as E
import hpt

RNG = np.random.default_rng(20240613)


def decode(data: bytes) -> np.ndarray:
    return hpt.decode_exr(data)


def expected_rgb(planes: dict) -> np.ndarray:
    h, w = next(iter(planes.values())).shape
    out = np.zeros((h, w, 3), np.float32)
    for k, name in enumerate("RGB"):
        if name in planes:
            out[..., k] = planes[name].astype(np.float32)
    if not any(n in planes for n in "RGB"):
        out[:] = planes["Y"].astype(np.float32)[..., None]
    return out


def same(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# --------------------------------------------------------- hand-assembled
def test_hand_assembled_uncompressed_file():
    """3x2 pixels, channels B, G, R as HALF written out field by field: the
    header attributes, one offset per scanline, then per line the B, G and R
    values of x = 0..2 little-endian."""
    def attr(name, kind, value):
        return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value

    ch = b"".join(n + b"\0" + struct.pack("<i", 1) + b"\0\0\0\0" + struct.pack("<ii", 1, 1) for n in (b"B", b"G", b"R"))
    head = (b"\x76\x2f\x31\x01" + struct.pack("<I", 2)
            + attr("channels", "chlist", ch + b"\0")
            + attr("compression", "compression", b"\x00")
            + attr("dataWindow", "box2i", struct.pack("<4i", 10, 20, 12, 21))
            + attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, 99, 99))
            + attr("lineOrder", "lineOrder", b"\x00")
            + attr("pixelAspectRatio", "float", struct.pack("<f", 1))
            + attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
            + attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    # half bits: 1.0 = 0x3C00, 2.0 = 0x4000, 0.5 = 0x3800, -1.5 = 0xBE00, 65504 = 0x7BFF, 2^-24 = 0x0001
    lines = [
        {"R": [0x3C00, 0x4000, 0x3800], "G": [0x0000, 0xBE00, 0x7BFF], "B": [0x0001, 0x3C00, 0x4000]},
        {"R": [0x3800, 0x3800, 0x3800], "G": [0x3C00, 0x3C00, 0x3C00], "B": [0x0000, 0x8000, 0x4000]},
    ]
    chunks = [struct.pack("<ii", 20 + y, 18) + b"".join(struct.pack("<3H", *ln[c]) for c in "BGR")
              for y, ln in enumerate(lines)]
    table = struct.pack("<QQ", len(head) + 16, len(head) + 16 + len(chunks[0]))
    img = decode(head + table + b"".join(chunks))
    want = np.array([[[1.0, 0.0, 2.0 ** -24], [2.0, -1.5, 1.0], [0.5, 65504.0, 2.0]],
                     [[0.5, 1.0, 0.0], [0.5, 1.0, -0.0], [0.5, 1.0, 2.0]]], np.float32)
    assert same(img, want)


def test_hand_assembled_piz_chunk():
    """One scanline of 100 HALF luminance values, 99 x 1.0 then 2.0, PIZ-coded
    by hand: the bitmap holds words 0x3C00 (byte 1920 bit 0) and 0x4000 (byte
    2048 bit 0), so the ranks are 0 -> 0, 0x3C00 -> 1, 0x4000 -> 2; one row
    takes no wavelet step; the Huffman table over symbols 1..3 (3 = run
    symbol) has lengths 1, 2, 2, so the canonical codes are 2 -> 00, 3 -> 01,
    1 -> 1; the stream is code(1), run symbol, 98 in 8 bits, code(2):
    1 01 01100010 00 = 0xAC 0x40, 13 bits."""
    bitmap = bytearray(129)
    bitmap[0] = 1
    bitmap[128] = 1
    table = bytes([0b00000100, 0b00100000, 0b10000000])  # 6-bit lengths 1, 2, 2
    huf = struct.pack("<5I", 1, 3, len(table), 13, 0) + table + bytes([0xAC, 0x40])
    body = struct.pack("<HH", 1920, 2048) + bytes(bitmap) + struct.pack("<i", len(huf)) + huf
    assert len(body) < 200
    head = E.header([("Y", E.HALF)], E.PIZ, (0, 0, 99, 0))
    data = head + struct.pack("<Q", len(head) + 8) + struct.pack("<ii", 0, len(body)) + body
    img = decode(data)
    want = np.ones((1, 100, 3), np.float32)
    want[0, 99] = 2.0
    assert same(img, want)


def test_every_half_bit_pattern_converts_like_numpy():
    bits = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    planes = {"Y": bits.view(np.float16)}
    img = decode(E.write_exr(planes, E.NONE))
    want = bits.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(img[..., 0]), nan)
    assert same(np.where(nan, 0, img[..., 0]).astype(np.float32), np.where(nan, 0, want).astype(np.float32))


# ------------------------------------------------------------ round trips
def _smooth(h, w, dtype, seed):
    r = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    v = 2.0 + np.sin(x * 0.21 + r.uniform(0, 6)) * np.cos(y * 0.17) + 0.05 * r.standard_normal((h, w))
    if dtype == np.uint32:
        return (v * 1000).astype(np.uint32)
    return (v * r.uniform(0.5, 40)).astype(dtype)


def _noise(h, w, dtype, seed):
    r = np.random.default_rng(seed)
    if dtype == np.float16:
        b = r.integers(0, 0x7C00, (h, w), dtype=np.uint16) | (r.integers(0, 2, (h, w), dtype=np.uint16) << 15)
        return b.view(np.float16)
    if dtype == np.float32:
        return (r.standard_normal((h, w)) * 10.0 ** r.uniform(-3, 3, (h, w))).astype(np.float32)
    return r.integers(0, 2 ** 32, (h, w), dtype=np.uint32)


SIZES = [(1, 1), (5, 7), (17, 33), (40, 64), (70, 3), (33, 16)]
COMPRESSIONS = [E.NONE, E.RLE, E.ZIPS, E.ZIP, E.PIZ]


@pytest.mark.parametrize("compression", COMPRESSIONS, ids=["none", "rle", "zips", "zip", "piz"])
@pytest.mark.parametrize("dtype", [np.float16, np.float32, np.uint32], ids=["half", "float", "uint"])
def test_round_trip(compression, dtype):
    for k, (h, w) in enumerate(SIZES):
        for make in (_smooth, _noise):
            planes = {c: make(h, w, dtype, 100 * k + i) for i, c in enumerate("RGB")}
            data = E.write_exr(planes, compression, x0=-3 + k, y0=5 - 2 * k)
            assert same(decode(data), expected_rgb(planes)), (h, w, make.__name__)


@pytest.mark.parametrize("compression", COMPRESSIONS, ids=["none", "rle", "zips", "zip", "piz"])
def test_mixed_channel_types_and_extra_channels(compression):
    h, w = 37, 29
    planes = {"A": _smooth(h, w, np.float16, 1), "R": _smooth(h, w, np.float16, 2),
              "G": _noise(h, w, np.float32, 3), "B": _smooth(h, w, np.uint32, 4), "Z": _noise(h, w, np.float32, 5)}
    assert same(decode(E.write_exr(planes, compression)), expected_rgb(planes))


def test_piz_exercises_both_wavelet_widths_and_raw_chunks():
    """Smooth data keeps the largest rank below 2^14 (14-bit lifting); 600
    columns of random halves exceed it (modular 16-bit lifting); a tiny noisy
    chunk does not shrink and is stored raw."""
    stats = {}
    smooth = {c: _smooth(64, 600, np.float16, i) for i, c in enumerate("RGB")}
    assert same(decode(E.write_exr(smooth, E.PIZ, stats=stats)), expected_rgb(smooth))
    assert max(stats["max_value"]) < 1 << 14
    stats = {}
    noisy = {c: _noise(40, 600, np.float16, 10 + i) for i, c in enumerate("RGB")}
    assert same(decode(E.write_exr(noisy, E.PIZ, stats=stats)), expected_rgb(noisy))
    assert max(stats["max_value"]) >= 1 << 14
    stats = {}
    tiny = {c: _noise(2, 2, np.float32, 20 + i) for i, c in enumerate("RGB")}
    assert same(decode(E.write_exr(tiny, E.PIZ, stats=stats)), expected_rgb(tiny))
    assert stats.get("raw_chunks", 0) >= 1


@pytest.mark.parametrize("compression", [E.ZIP, E.PIZ], ids=["zip", "piz"])
def test_decreasing_line_order(compression):
    planes = {c: _smooth(70, 20, np.float16, 30 + i) for i, c in enumerate("RGB")}
    assert same(decode(E.write_exr(planes, compression, line_order=1)), expected_rgb(planes))


def test_luminance_only_and_missing_channels():
    y = _smooth(9, 11, np.float16, 40)
    img = decode(E.write_exr({"Y": y}, E.ZIP))
    assert same(img, np.repeat(y.astype(np.float32)[..., None], 3, axis=2))
    rg = {"R": _smooth(9, 11, np.float16, 41), "G": _smooth(9, 11, np.float16, 42), "Y": y}
    img = decode(E.write_exr(rg, E.PIZ))
    assert same(img, expected_rgb({"R": rg["R"], "G": rg["G"]}))
    assert not img[..., 2].any()


def test_runs_in_piz_and_rle():
    planes = {c: np.full((50, 300), v, np.float16) for c, v in zip("RGB", (0.0, 1.0, 3.5))}
    planes["G"][10:12, 100:250] = 7.0
    for comp in (E.RLE, E.PIZ, E.ZIP):
        data = E.write_exr(planes, comp)
        assert len(data) < 50 * 300 * 6 // 4
        assert same(decode(data), expected_rgb(planes))


def test_read_env_reads_exr_files(tmp_path):
    import app

    planes = {c: _smooth(16, 32, np.float32, 50 + i) for i, c in enumerate("RGB")}
    path = os.path.join(tmp_path, "sky.exr")
    with open(path, "wb") as f:
        f.write(E.write_exr(planes, E.PIZ))
    assert same(app.read_env(path), expected_rgb(planes))
    assert same(hpt.read_exr(path), expected_rgb(planes))


# ------------------------------------------------------------------ errors
def _status(data: bytes):
    try:
        decode(data)
    except hpt.HptError as e:
        return e.status, str(e)
    return 0, ""


def test_unsupported_files_are_rejected_with_a_message():
    planes = {c: _smooth(8, 8, np.float16, 60 + i) for i, c in enumerate("RGB")}
    good = E.write_exr(planes, E.ZIP)
    tiled = good[:4] + struct.pack("<I", 2 | 0x200) + good[8:]
    assert _status(tiled)[0] == 5 and "tiled" in _status(tiled)[1]
    multi = good[:4] + struct.pack("<I", 2 | 0x1000) + good[8:]
    assert _status(multi)[0] == 5
    pxr = E.header([("R", E.HALF)], E.PXR24, (0, 0, 7, 7)) + struct.pack("<Q", 0)
    st, msg = _status(pxr)
    assert st == 5 and "PXR24" in msg
    sub = E.header([("R", E.HALF)], E.NONE, (0, 0, 7, 7), sampling=(2, 2)) + struct.pack("<8Q", *[0] * 8)
    assert _status(sub)[0] == 5 and "subsampled" in _status(sub)[1]
    depth = E.write_exr({"Z": _noise(4, 4, np.float32, 1)}, E.NONE)
    st, msg = _status(depth)
    assert st == 5 and "Z" in msg


def test_malformed_files_fail_cleanly():
    planes = {c: _smooth(40, 37, np.float16, 70 + i) for i, c in enumerate("RGB")}
    assert _status(b"")[0] == 1
    assert _status(b"\x76\x2f\x31\x02" + bytes(100))[0] == 1
    for comp in COMPRESSIONS:
        good = E.write_exr(planes, comp)
        for cut in (10, len(good) // 2, len(good) - 1):
            assert _status(good[:cut])[0] == 1, (comp, cut)
    # Random byte flips: either a clean error or some image, never a crash.
    r = np.random.default_rng(7)
    for comp in COMPRESSIONS:
        good = bytearray(E.write_exr(planes, comp))
        for _ in range(60):
            bad = bytearray(good)
            for pos in r.integers(0, len(bad), size=r.integers(1, 4)):
                bad[pos] = int(r.integers(0, 256))
            st, _ = _status(bytes(bad))
            assert st in (0, 1, 5)


def test_zip_chunk_must_inflate_to_its_size():
    planes = {"R": _smooth(16, 16, np.float16, 80)}
    head = E.header([("R", E.HALF)], E.ZIP, (0, 0, 15, 15))
    body = zlib.compress(b"\0" * 100)  # 100 bytes, not 512
    data = head + struct.pack("<Q", len(head) + 8) + struct.pack("<ii", 0, len(body)) + body
    st, msg = _status(data)
    assert st == 1 and "inflate" in msg
    assert same(decode(E.write_exr(planes, E.ZIP)), expected_rgb(planes))
