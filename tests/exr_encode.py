"""OpenEXR writer for the decoder's tests (test infrastructure only).

Writes single-part scanline files: magic 76 2f 31 01, version 2, the
attribute header (channels, compression, dataWindow, displayWindow,
lineOrder, pixelAspectRatio, screenWindowCenter, screenWindowWidth), the
offset table and the chunks, with the encoding steps of the NONE, RLE, ZIPS,
ZIP and PIZ compressors:

- RLE/ZIP: bytes split into even and odd halves, then d[i] = t[i] - t[i-1]
  + 128 (mod 256), then runs (RLE) or a zlib stream (ZIP).
- PIZ: 16-bit words in channel planes; a bitmap of the words present (word 0
  implied); the ranks of the words; a 2D Haar lifting wavelet per plane
  (14-bit when the largest rank is below 2^14, modular 16-bit otherwise);
  a Huffman code of the result with a run-length symbol and a canonical
  code table whose longest codes are numbered from 0.

A chunk whose encoding is not smaller than its raw bytes is stored raw, as
the OpenEXR writer does.  Everything here is small-image Python."""
from __future__ import annotations

import heapq
import struct
import zlib

import numpy as np

UINT, HALF, FLOAT = 0, 1, 2
NONE, RLE, ZIPS, ZIP, PIZ, PXR24 = 0, 1, 2, 3, 4, 5
LINES = {NONE: 1, RLE: 1, ZIPS: 1, ZIP: 16, PIZ: 32}
DTYPES = {UINT: "<u4", HALF: "<f2", FLOAT: "<f4"}


def attribute(name: str, kind: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def chlist(channels) -> bytes:
    out = b""
    for name, kind in sorted(channels):
        out += name.encode() + b"\0" + struct.pack("<iB3xii", kind, 0, 1, 1)
    return out + b"\0"


def header(channels, compression, window, line_order=0, flags=0, sampling=None) -> bytes:
    x0, y0, x1, y1 = window
    ch = chlist(channels)
    if sampling is not None:  # (x, y) sampling written for every channel
        ch = b""
        for name, kind in sorted(channels):
            ch += name.encode() + b"\0" + struct.pack("<iB3xii", kind, 0, *sampling)
        ch += b"\0"
    attrs = [
        attribute("channels", "chlist", ch),
        attribute("compression", "compression", bytes([compression])),
        attribute("dataWindow", "box2i", struct.pack("<4i", x0, y0, x1, y1)),
        attribute("displayWindow", "box2i", struct.pack("<4i", x0, y0, x1, y1)),
        attribute("lineOrder", "lineOrder", bytes([line_order])),
        attribute("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        attribute("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0)),
        attribute("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ]
    return b"\x76\x2f\x31\x01" + struct.pack("<I", 2 | flags) + b"".join(attrs) + b"\0"


# --------------------------------------------------------------- RLE / ZIP
def _predict(raw: bytes) -> bytes:
    a = np.frombuffer(raw, np.uint8)
    t = np.concatenate([a[0::2], a[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _rle(data: bytes) -> bytes:
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i + 1
        while j < n and data[j] == data[i] and j - i < 128:
            j += 1
        if j - i >= 3:
            out += struct.pack("<bB", j - i - 1, data[i])
            i = j
            continue
        j = i
        while j < n and j - i < 127 and not (j + 2 < n and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        out += struct.pack("<b", -(j - i)) + data[i:j]
        i = j
    return bytes(out)


# --------------------------------------------------------------------- PIZ
def _lift14(a, b):
    """Signed 16-bit pair: mean floor((a + b) / 2) and difference a - b."""
    sa = a - 65536 if a >= 32768 else a
    sb = b - 65536 if b >= 32768 else b
    return ((sa + sb) >> 1) & 0xFFFF, (sa - sb) & 0xFFFF


def _lift16(a, b):
    """Modular pair: a offset by 2^15; the mean moves by 2^15 when the
    difference is negative."""
    ao = (a + 0x8000) & 0xFFFF
    m, d = (ao + b) >> 1, ao - b
    if d < 0:
        m = (m + 0x8000) & 0xFFFF
    return m, d & 0xFFFF


def wav2_encode(buf, start, nx, ox, ny, oy, mx):
    """The forward 2D transform of one plane of `buf` (a list of words; word
    (x, y) at start + y * oy + x * ox): pair distances 1, 2, 4, ... while
    twice the distance fits min(nx, ny); at each level rows, then columns; a
    leftover column or row (bit s of nx or ny set) takes the 1D pair."""
    lift = _lift14 if mx < (1 << 14) else _lift16

    def idx(x, y):
        return start + y * oy + x * ox

    s = 1
    while 2 * s <= min(nx, ny):
        step = 2 * s
        y = 0
        while y + step <= ny:
            x = 0
            while x + step <= nx:
                q00, q01, q10, q11 = idx(x, y), idx(x + s, y), idx(x, y + s), idx(x + s, y + s)
                t0, t1 = lift(buf[q00], buf[q01])      # row y
                u0, u1 = lift(buf[q10], buf[q11])      # row y + s
                buf[q00], buf[q10] = lift(t0, u0)      # column x
                buf[q01], buf[q11] = lift(t1, u1)      # column x + s
                x += step
            if nx & s:
                q0, q1 = idx(x, y), idx(x, y + s)
                buf[q0], buf[q1] = lift(buf[q0], buf[q1])
            y += step
        if ny & s:
            x = 0
            while x + step <= nx:
                q0, q1 = idx(x, y), idx(x + s, y)
                buf[q0], buf[q1] = lift(buf[q0], buf[q1])
                x += step
        s *= 2


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0
        self.total = 0

    def put(self, value: int, nbits: int):
        for k in range(nbits - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> k) & 1)
            self.n += 1
            self.total += 1
            if self.n == 8:
                self.out.append(self.acc)
                self.acc = self.n = 0

    def flush(self) -> bytes:
        if self.n:
            self.out.append((self.acc << (8 - self.n)) & 0xFF)
            self.acc = self.n = 0
        return bytes(self.out)


def huffman_lengths(freq: dict) -> dict:
    heap = [(f, s, [s]) for s, f in sorted(freq.items())]
    heapq.heapify(heap)
    length = {s: 0 for s in freq}
    while len(heap) > 1:
        f1, k1, a = heapq.heappop(heap)
        f2, k2, b = heapq.heappop(heap)
        for s in a + b:
            length[s] += 1
        heapq.heappush(heap, (f1 + f2, min(k1, k2), a + b))
    assert max(length.values()) <= 58
    return length


def canonical_codes(length: dict) -> dict:
    count = [0] * 59
    for l in length.values():
        count[l] += 1
    start = [0] * 59
    c = 0
    for l in range(58, 0, -1):
        start[l] = c
        c = (c + count[l]) >> 1
    codes = {}
    for s in sorted(length):
        l = length[s]
        codes[s] = (start[l], l)
        start[l] += 1
    return codes


def huf_compress(words) -> bytes:
    if len(words) == 0:
        return b""
    freq = {}
    for w in words:
        freq[w] = freq.get(w, 0) + 1
    im, iM = min(freq), max(freq) + 1
    freq[iM] = 1  # run-length symbol
    codes = canonical_codes(huffman_lengths(freq))
    table = _Bits()
    i = im
    while i <= iM:
        l = codes[i][1] if i in codes else 0
        if l == 0:
            run = 1
            while i + run <= iM and (i + run) not in codes and run < 255 + 6:
                run += 1
            if run >= 6:
                table.put(63, 6)
                table.put(run - 6, 8)
                i += run
                continue
            if run >= 2:
                table.put(59 + run - 2, 6)
                i += run
                continue
        table.put(l, 6)
        i += 1
    table_bytes = table.flush()
    data = _Bits()

    def send(s, cs):
        cl, rl = codes[s][1], codes[iM][1]
        if cl + rl + 8 < cl * cs:
            data.put(*codes[s])
            data.put(*codes[iM])
            data.put(cs, 8)
        else:
            for _ in range(cs + 1):
                data.put(*codes[s])

    s, cs = words[0], 0
    for w in words[1:]:
        if w == s and cs < 255:
            cs += 1
        else:
            send(s, cs)
            cs = 0
        s = w
    send(s, cs)
    nbits = data.total
    return struct.pack("<5I", im, iM, len(table_bytes), nbits, 0) + table_bytes + data.flush()


def piz_compress(raw: bytes, width: int, lines: int, channels) -> tuple:
    """Returns (bytes, largest rank) for one chunk's raw scanline bytes."""
    chans = sorted(channels)
    sizes = [2 if kind == HALF else 4 for _, kind in chans]
    planes = [[] for _ in chans]
    pos = 0
    for _ in range(lines):
        for c, sz in enumerate(sizes):
            n = width * sz
            planes[c].extend(np.frombuffer(raw[pos:pos + n], "<u2").tolist())
            pos += n
    words = [w for p in planes for w in p]
    bitmap = bytearray(8192)
    for w in set(words):
        bitmap[w >> 3] |= 1 << (w & 7)
    bitmap[0] &= 0xFE
    nz = [i for i in range(8192) if bitmap[i]]
    lo, hi = (nz[0], nz[-1]) if nz else (8191, 0)
    lut, k = {}, 0
    for v in range(65536):
        if v == 0 or bitmap[v >> 3] & (1 << (v & 7)):
            lut[v] = k
            k += 1
    max_value = k - 1
    words = [lut[w] for w in words]
    at = 0
    for c, sz in enumerate(sizes):
        n = len(planes[c])
        for j in range(sz // 2):
            wav2_encode(words, at + j, width, sz // 2, lines, width * sz // 2, max_value)
        at += n
    huf = huf_compress(words)
    out = struct.pack("<HH", lo, hi)
    if lo <= hi:
        out += bytes(bitmap[lo:hi + 1])
    return out + struct.pack("<i", len(huf)) + huf, max_value


# ------------------------------------------------------------------- files
def write_exr(planes: dict, compression: int, x0: int = 0, y0: int = 0, line_order: int = 0,
              stats: dict | None = None) -> bytes:
    """`planes`: channel name -> (H, W) array of float16, float32 or uint32."""
    kinds = {np.dtype(np.float16): HALF, np.dtype(np.float32): FLOAT, np.dtype(np.uint32): UINT}
    channels = [(name, kinds[a.dtype]) for name, a in planes.items()]
    h, w = next(iter(planes.values())).shape
    window = (x0, y0, x0 + w - 1, y0 + h - 1)
    head = header(channels, compression, window, line_order)
    lpc = LINES[compression]
    chunks = []
    for y in range(0, h, lpc):
        n = min(lpc, h - y)
        raw = b"".join(planes[name][y + l].astype(DTYPES[kind]).tobytes()
                       for l in range(n) for name, kind in sorted(channels))
        if compression == NONE:
            body = raw
        elif compression == RLE:
            body = _rle(_predict(raw))
        elif compression in (ZIPS, ZIP):
            body = zlib.compress(_predict(raw))
        else:
            body, max_value = piz_compress(raw, w, n, channels)
            if stats is not None:
                stats.setdefault("max_value", []).append(max_value)
        if len(body) >= len(raw):
            body = raw
            if stats is not None:
                stats["raw_chunks"] = stats.get("raw_chunks", 0) + 1
        chunks.append(struct.pack("<ii", y0 + y, len(body)) + body)
    order = list(range(len(chunks)))
    if line_order == 1:
        order.reverse()
    base = len(head) + 8 * len(chunks)
    offsets = [0] * len(chunks)
    blob = b""
    for i in order:
        offsets[i] = base + len(blob)
        blob += chunks[i]
    return head + struct.pack(f"<{len(chunks)}Q", *offsets) + blob
