// exr.cpp — reader of single-part scanline OpenEXR files, the format of the
// HDRI environment maps the reference's configurations name (conf/*.yaml
// `hdri_file`, handed to HalaRenderer::set_envmap at
// src/bin/hala-pathtracer.rs:202-203).
//
// The file (OpenEXR layout, version 2):
//   magic 76 2f 31 01; u32 version 2 | flags (0x200 tiled, 0x400 long
//   names, 0x800 deep, 0x1000 multi-part)
//   header   attributes `name\0type\0` i32 size, value; a 0 byte ends it
//   offsets  one u64 file offset per chunk
//   chunk    i32 y of its first line, i32 byte count, bytes
// A chunk covers `lines` scanlines (NONE, RLE, ZIPS 1; ZIP 16; PIZ 32).  Its
// unpacked bytes hold, per scanline, every channel in header (name) order,
// `width` little-endian values each (UINT 4, HALF 2, FLOAT 4 bytes).  A chunk
// whose byte count is not below its unpacked size is stored unpacked.
//
// Compressions:
//   RLE  a signed byte n < 0 copies -n literal bytes, n >= 0 repeats the next
//        byte n + 1 times; then the byte reordering of ZIP
//   ZIP  a zlib stream; then t[i] += t[i-1] - 128 (mod 256) and the halves
//        interleaved (first half to the even bytes, second to the odd)
//   PIZ  the chunk as 16-bit words, one plane per channel; a bitmap of the
//        words present and the table that maps their ranks back; a Huffman
//        code of the ranks with a run-length symbol; a lossless 2D Haar
//        wavelet per plane (14-bit or modular 16-bit lifting)
// Rejected with a message: tiled, deep and multi-part files, subsampled
// channels, and the PXR24, B44, B44A, DWAA and DWAB compressions.
#include "exr.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace hpt {
namespace {

struct Bad : std::runtime_error {
  ExrResult kind;
  Bad(ExrResult k, const std::string& m) : std::runtime_error(m), kind(k) {}
};

void need(bool c, const std::string& m) {
  if (!c) throw Bad(kExrMalformed, m);
}
void supported(bool c, const std::string& m) {
  if (!c) throw Bad(kExrUnsupported, m);
}

uint32_t rd32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

struct In {
  const uint8_t* p;
  size_t n;
  size_t pos;
  In(const uint8_t* d, size_t s, size_t at = 0) : p(d), n(s), pos(at) {}
  size_t left() const { return n - pos; }
  uint8_t u8() {
    need(left() >= 1, "file truncated");
    return p[pos++];
  }
  uint16_t u16() {
    need(left() >= 2, "file truncated");
    uint16_t v = (uint16_t)(p[pos] | (p[pos + 1] << 8));
    pos += 2;
    return v;
  }
  uint32_t u32() {
    need(left() >= 4, "file truncated");
    uint32_t v = rd32(p + pos);
    pos += 4;
    return v;
  }
  int32_t i32() { return (int32_t)u32(); }
  uint64_t u64() {
    uint64_t lo = u32();
    return lo | ((uint64_t)u32() << 32);
  }
  std::string cstr(size_t max_len) {
    size_t s = pos;
    while (pos < n && p[pos]) ++pos;
    need(pos < n, "header truncated");
    std::string r((const char*)p + s, pos - s);
    ++pos;
    need(r.size() <= max_len, "attribute or channel name too long");
    return r;
  }
};

enum { kUint = 0, kHalf = 1, kFloat = 2 };
enum { kNone = 0, kRle = 1, kZips = 2, kZip = 3, kPiz = 4 };
const char* const kCompressionNames[] = {"NONE", "RLE", "ZIPS", "ZIP", "PIZ", "PXR24", "B44", "B44A", "DWAA", "DWAB"};

struct Channel {
  std::string name;
  int type;
  uint32_t bytes;
};

struct Header {
  std::vector<Channel> channels;
  int compression = -1;
  bool have_window = false;
  int32_t x0 = 0, y0 = 0, x1 = -1, y1 = -1;
};

uint32_t lines_per_chunk(int compression) {
  return compression == kZip ? 16u : compression == kPiz ? 32u : 1u;
}

Header read_header(In& in) {
  need(in.n >= 8 && in.p[0] == 0x76 && in.p[1] == 0x2f && in.p[2] == 0x31 && in.p[3] == 0x01,
       "not an OpenEXR file (bad magic number)");
  in.pos = 4;
  const uint32_t version = in.u32();
  supported((version & 0xFF) == 2, "OpenEXR file version " + std::to_string(version & 0xFF) + " (only 2 is read)");
  supported(!(version & 0x200), "tiled OpenEXR files are not supported (scanline files only)");
  supported(!(version & 0x800), "deep OpenEXR files are not supported");
  supported(!(version & 0x1000), "multi-part OpenEXR files are not supported");
  need((version & ~0x1EFFu) == 0, "unknown OpenEXR version flags");
  const size_t max_name = (version & 0x400) ? 255 : 31;

  Header h;
  for (;;) {
    const std::string name = in.cstr(max_name);
    if (name.empty()) break;
    const std::string type = in.cstr(max_name);
    const int32_t size = in.i32();
    need(size >= 0 && (size_t)size <= in.left(), "attribute " + name + " runs past the end of the file");
    In v(in.p, in.pos + (size_t)size, in.pos);
    if (name == "channels" && type == "chlist") {
      for (;;) {
        const std::string cname = v.cstr(max_name);
        if (cname.empty()) break;
        const int32_t t = v.i32();
        v.u8();  // pLinear
        v.u8();
        v.u8();
        v.u8();  // reserved
        const int32_t xs = v.i32(), ys = v.i32();
        need(t >= kUint && t <= kFloat, "channel " + cname + " has an unknown pixel type");
        supported(xs == 1 && ys == 1, "channel " + cname + " is subsampled (only x/y sampling 1 is read)");
        h.channels.push_back(Channel{cname, t, t == kHalf ? 2u : 4u});
      }
    } else if (name == "compression" && type == "compression") {
      need(size == 1, "bad compression attribute");
      h.compression = v.u8();
    } else if (name == "dataWindow" && type == "box2i") {
      need(size == 16, "bad dataWindow attribute");
      h.x0 = v.i32();
      h.y0 = v.i32();
      h.x1 = v.i32();
      h.y1 = v.i32();
      h.have_window = true;
    } else if (name == "type" && type == "string") {
      const std::string s((const char*)in.p + in.pos, (size_t)size);
      supported(s == "scanlineimage", "OpenEXR part type " + s + " is not supported (scanlineimage only)");
    }
    in.pos += (size_t)size;
  }
  need(!h.channels.empty(), "no channels attribute");
  need(h.compression >= 0, "no compression attribute");
  need(h.have_window, "no dataWindow attribute");
  supported(h.compression <= kPiz,
            std::string("compression ") +
                (h.compression < 10 ? kCompressionNames[h.compression] : std::to_string(h.compression).c_str()) +
                " is not supported (NONE, RLE, ZIPS, ZIP and PIZ are)");
  need(h.x1 >= h.x0 && h.y1 >= h.y0, "empty data window");
  const int64_t w = (int64_t)h.x1 - h.x0 + 1, ht = (int64_t)h.y1 - h.y0 + 1;
  need(w <= (1 << 20) && ht <= (1 << 20) && w * ht <= (1ll << 28), "data window too large");
  return h;
}

float half_to_float(uint16_t v) {
  const uint32_t s = (uint32_t)(v >> 15) << 31, e = (v >> 10) & 31u, m = v & 1023u;
  if (e == 0) {
    const float f = std::ldexp((float)m, -24);
    return s ? -f : f;
  }
  const uint32_t bits = s | ((e == 31 ? 255u : e + 112u) << 23) | (m << 13);
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// ------------------------------------------------------------- RLE and ZIP
void rle_unpack(const uint8_t* src, size_t n, uint8_t* out, size_t out_n) {
  size_t i = 0, o = 0;
  while (i < n) {
    const int8_t c = (int8_t)src[i++];
    if (c < 0) {
      const size_t k = (size_t)(-(int)c);
      need(k <= n - i && k <= out_n - o, "RLE chunk overruns");
      std::memcpy(out + o, src + i, k);
      i += k;
      o += k;
    } else {
      const size_t k = (size_t)c + 1;
      need(i < n && k <= out_n - o, "RLE chunk overruns");
      std::memset(out + o, src[i++], k);
      o += k;
    }
  }
  need(o == out_n, "RLE chunk decodes to the wrong size");
}

// The byte predictor and half-interleave that RLE and ZIP chunks undo.
void unpredict_interleave(uint8_t* t, uint8_t* out, size_t n) {
  for (size_t i = 1; i < n; ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
  const uint8_t* t1 = t;
  const uint8_t* t2 = t + (n + 1) / 2;
  for (size_t i = 0; i < n; ++i) out[i] = (i & 1) ? *t2++ : *t1++;
}

void zip_unpack(const uint8_t* src, size_t n, uint8_t* out, size_t out_n) {
  uLongf len = (uLongf)out_n;
  const int r = uncompress(out, &len, src, (uLong)n);
  need(r == Z_OK && len == out_n, "ZIP chunk does not inflate to its unpacked size");
}

// --------------------------------------------------------------- PIZ Huffman
constexpr uint32_t kEncSize = 65537;  // 2^16 words + the run-length symbol
constexpr uint32_t kShortZeroRun = 59, kLongZeroRun = 63;
constexpr uint32_t kShortestLongRun = 2 + kLongZeroRun - kShortZeroRun;
constexpr int kMaxCodeLen = 58;
constexpr int kFastBits = 12;

// Decodes `out_n` words.  The stream: u32 im, iM, table bytes, bit count, 0;
// code lengths of symbols im..iM in 6-bit fields (59..62: 2..5 zero lengths,
// 63 + 8 bits: 6..261 zero lengths); then the codes, MSB first.  Symbol iM
// is the run-length symbol: 8 more bits repeat the previous word that often.
// Codes are canonical with the longest codes numbered from 0: walking
// lengths from 58 down to 1, codes of length l start at c and c becomes
// (c + count[l]) / 2.
void huf_unpack(const uint8_t* src, size_t n, uint16_t* out, size_t out_n) {
  if (n == 0) {
    need(out_n == 0, "PIZ chunk has no Huffman stream");
    return;
  }
  need(n >= 20, "PIZ Huffman header truncated");
  const uint32_t im = rd32(src), iM = rd32(src + 4), nbits = rd32(src + 12);
  need(im <= iM && iM < kEncSize, "PIZ Huffman table range out of bounds");

  std::vector<uint8_t> len(kEncSize, 0);
  size_t p = 20;
  uint64_t acc = 0;
  int lc = 0;
  auto get = [&](int nb) -> uint32_t {
    while (lc < nb) {
      need(p < n, "PIZ Huffman table truncated");
      acc = (acc << 8) | src[p++];
      lc += 8;
    }
    lc -= nb;
    return (uint32_t)((acc >> lc) & ((1u << nb) - 1));
  };
  for (uint32_t i = im; i <= iM; ++i) {
    const uint32_t l = get(6);
    uint32_t run = 0;
    if (l == kLongZeroRun) {
      run = get(8) + kShortestLongRun;
    } else if (l >= kShortZeroRun) {
      run = l - kShortZeroRun + 2;
    } else {
      len[i] = (uint8_t)l;
      continue;
    }
    need(run <= iM + 1 - i, "PIZ Huffman table too long");
    i += run - 1;
  }
  need((uint64_t)nbits <= 8ull * (n - p), "PIZ Huffman bit count exceeds the chunk");

  uint64_t count[kMaxCodeLen + 1] = {}, start[kMaxCodeLen + 1] = {};
  for (uint32_t i = im; i <= iM; ++i) count[len[i]]++;
  uint64_t c = 0;
  for (int l = kMaxCodeLen; l > 0; --l) {
    start[l] = c;
    c = (c + count[l]) >> 1;
    need(start[l] + count[l] <= (1ull << l), "PIZ Huffman code lengths overflow");
  }
  uint32_t first[kMaxCodeLen + 2] = {};
  for (int l = 1; l <= kMaxCodeLen; ++l) first[l + 1] = first[l] + (uint32_t)count[l];
  std::vector<uint32_t> syms(first[kMaxCodeLen + 1]);
  {
    uint32_t fill[kMaxCodeLen + 1];
    std::memcpy(fill, first, sizeof(fill));
    for (uint32_t i = im; i <= iM; ++i)
      if (len[i]) syms[fill[len[i]]++] = i;
  }
  // Codes of up to kFastBits bits decode through one table look-up.
  std::vector<uint32_t> fast(1u << kFastBits, 0);  // (symbol << 8) | length; 0 = longer code
  for (int l = 1; l <= kFastBits; ++l)
    for (uint32_t k = 0; k < count[l]; ++k) {
      const uint32_t base = (uint32_t)(start[l] + k) << (kFastBits - l);
      for (uint32_t s = 0; s < (1u << (kFastBits - l)); ++s) {
        need(fast[base + s] == 0, "PIZ Huffman table is not a prefix code");
        fast[base + s] = (syms[first[l] + k] << 8) | (uint32_t)l;
      }
    }

  const uint8_t* d = src + p;
  const size_t dbytes = n - p;
  const uint64_t total = nbits;
  auto byte_at = [&](uint64_t b) -> uint32_t { return b < dbytes ? d[b] : 0u; };
  auto bits_at = [&](uint64_t q, int k) -> uint32_t {  // k <= 12 bits from bit q, MSB first
    const uint64_t b = q >> 3;
    const uint32_t w = (byte_at(b) << 16) | (byte_at(b + 1) << 8) | byte_at(b + 2);
    return (w >> (24 - k - (int)(q & 7))) & ((1u << k) - 1);
  };
  uint64_t pos = 0;
  size_t o = 0;
  while (pos < total) {
    uint32_t sym = 0;
    const uint32_t e = fast[bits_at(pos, kFastBits)];
    if (e && (e & 0xFF) <= total - pos) {
      sym = e >> 8;
      pos += e & 0xFF;
    } else {
      uint64_t code = 0;
      int l = 0;
      bool found = false;
      while (pos < total && l < kMaxCodeLen) {
        code = (code << 1) | bits_at(pos++, 1);
        ++l;
        if (count[l] && code >= start[l] && code - start[l] < count[l]) {
          sym = syms[first[l] + (uint32_t)(code - start[l])];
          found = true;
          break;
        }
      }
      need(found, "PIZ Huffman stream holds an invalid code");
    }
    if (sym == iM) {
      need(total - pos >= 8, "PIZ Huffman run length truncated");
      const uint32_t run = bits_at(pos, 8);
      pos += 8;
      need(o > 0, "PIZ Huffman run before the first word");
      need(run <= out_n - o, "PIZ Huffman stream decodes too many words");
      std::fill(out + o, out + o + run, out[o - 1]);
      o += run;
    } else {
      need(o < out_n, "PIZ Huffman stream decodes too many words");
      out[o++] = (uint16_t)sym;
    }
  }
  need(o == out_n, "PIZ Huffman stream decodes too few words");
}

// ------------------------------------------------------------- PIZ wavelet
// PIZ stores each channel plane as a 2D Haar transform built from one 1D
// lifting pair per two words: the mean m and difference d of (a, b).  Two
// variants exist and the file's largest word rank picks one:
//   ranks below 2^14: signed 16-bit, m = floor((a + b) / 2), d = a - b, so
//     a = m + ceil(d / 2) and b = a - d;
//   otherwise modular 16-bit: a is offset by 2^15, m = floor((a' + b) / 2),
//     d = a' - b, and m moves by 2^15 when d < 0, so b = m - (d >> 1) and
//     a = d + b - 2^15 (all mod 2^16).
// The forward transform runs its levels from the finest pair distance
// (1) to the coarsest (half the largest power of two <= min(nx, ny)); at a
// level with pair distance s it combines the words at (x, y), (x + s, y),
// (x, y + s), (x + s, y + s) for x, y on the grid of spacing 2s, rows
// first, then columns; a column (row) left over when nx (ny) has bit s set
// takes the one-dimensional pair.  Decoding undoes the levels from the
// coarsest, columns before rows.  Written from that definition; files
// decode bit for bit only with exactly these steps and this order.
struct LiftPair {
  bool small;  // the signed 14-bit variant
  void undo(uint16_t m, uint16_t d, uint16_t& a, uint16_t& b) const {
    if (small) {
      const int mi = (int16_t)m, di = (int16_t)d;
      const int ai = mi + ((di + (di & 1)) >> 1);
      a = (uint16_t)ai;
      b = (uint16_t)(ai - di);
    } else {
      const uint32_t bb = ((uint32_t)m - (uint32_t)(d >> 1)) & 0xFFFFu;
      b = (uint16_t)bb;
      a = (uint16_t)(((uint32_t)d + bb + 0x8000u) & 0xFFFFu);  // + 2^15 == - 2^15 (mod 2^16)
    }
  }
};

// Undoes the transform of one plane of nx x ny words; word (x, y) is at
// plane[y * row + x * col].  max_rank: the largest word rank of the chunk.
void piz_unwavelet(uint16_t* plane, int nx, int col, int ny, int row, uint16_t max_rank) {
  const LiftPair lift{max_rank < (1u << 14)};
  auto word = [&](int x, int y) -> uint16_t& { return plane[(ptrdiff_t)y * row + (ptrdiff_t)x * col]; };
  const int n = std::min(nx, ny);
  int coarse = 1;  // the largest power of two <= n
  while (coarse * 2 <= n) coarse *= 2;
  for (int s = coarse / 2; s >= 1; s /= 2) {
    const int step = 2 * s;
    int y = 0;
    for (; y + step <= ny; y += step) {
      int x = 0;
      for (; x + step <= nx; x += step) {
        uint16_t top0, top1, bot0, bot1;
        lift.undo(word(x, y), word(x, y + s), top0, bot0);          // columns x and x + s
        lift.undo(word(x + s, y), word(x + s, y + s), top1, bot1);
        lift.undo(top0, top1, word(x, y), word(x + s, y));            // rows y and y + s
        lift.undo(bot0, bot1, word(x, y + s), word(x + s, y + s));
      }
      if (nx & s) {  // the leftover column: its vertical pair
        uint16_t u, v;
        lift.undo(word(x, y), word(x, y + s), u, v);
        word(x, y) = u;
        word(x, y + s) = v;
      }
    }
    if (ny & s) {  // the leftover row: its horizontal pairs
      for (int x = 0; x + step <= nx; x += step) {
        uint16_t u, v;
        lift.undo(word(x, y), word(x + s, y), u, v);
        word(x, y) = u;
        word(x + s, y) = v;
      }
    }
  }
}

void piz_unpack(const uint8_t* src, size_t n, uint32_t width, uint32_t lines, const std::vector<Channel>& chans,
                uint8_t* out, size_t out_n) {
  const size_t words = out_n / 2;
  std::vector<uint16_t> tmp(words);
  In in(src, n);
  const uint16_t min_nz = in.u16(), max_nz = in.u16();
  need(max_nz < 8192, "PIZ bitmap range out of bounds");
  std::vector<uint8_t> bitmap(8192, 0);
  if (min_nz <= max_nz) {
    const size_t k = (size_t)max_nz - min_nz + 1;
    need(in.left() >= k, "PIZ bitmap truncated");
    std::memcpy(bitmap.data() + min_nz, in.p + in.pos, k);
    in.pos += k;
  }
  // Rank -> word: word 0 always has rank 0, then every word in the bitmap.
  std::vector<uint16_t> lut(65536, 0);
  uint32_t k = 0;
  for (uint32_t i = 0; i < 65536; ++i)
    if (i == 0 || (bitmap[i >> 3] & (1u << (i & 7)))) lut[k++] = (uint16_t)i;
  const uint16_t max_value = (uint16_t)(k - 1);
  const int32_t hlen = in.i32();
  need(hlen >= 0 && (size_t)hlen <= in.left(), "PIZ Huffman stream runs past the chunk");
  huf_unpack(in.p + in.pos, (size_t)hlen, tmp.data(), words);

  size_t at = 0;
  for (const Channel& c : chans) {
    const int sz = (int)c.bytes / 2;
    for (int j = 0; j < sz; ++j) piz_unwavelet(tmp.data() + at + j, (int)width, sz, (int)lines, (int)width * sz, max_value);
    at += (size_t)width * lines * sz;
  }
  for (uint16_t& w : tmp) w = lut[w];

  std::vector<size_t> cur(chans.size());
  at = 0;
  for (size_t c = 0; c < chans.size(); ++c) {
    cur[c] = at;
    at += (size_t)width * lines * (chans[c].bytes / 2);
  }
  uint8_t* q = out;
  for (uint32_t y = 0; y < lines; ++y)
    for (size_t c = 0; c < chans.size(); ++c) {
      const size_t nw = (size_t)width * (chans[c].bytes / 2);
      for (size_t i = 0; i < nw; ++i) {
        const uint16_t v = tmp[cur[c] + i];
        *q++ = (uint8_t)(v & 0xFF);
        *q++ = (uint8_t)(v >> 8);
      }
      cur[c] += nw;
    }
}

}  // namespace

ExrResult decode_exr(const uint8_t* data, size_t size, bool pixels, ExrImage& out, std::string& err) {
  try {
    need(data != nullptr, "null data");
    In in(data, size);
    const Header h = read_header(in);
    const uint32_t width = (uint32_t)(h.x1 - h.x0 + 1), height = (uint32_t)(h.y1 - h.y0 + 1);
    const uint32_t lpc = lines_per_chunk(h.compression);
    const uint32_t chunks = (height + lpc - 1) / lpc;
    std::vector<uint64_t> offsets(chunks);
    for (auto& o : offsets) o = in.u64();
    out.width = width;
    out.height = height;
    out.rgb.clear();
    if (!pixels) return kExrOk;

    int target[3] = {-1, -1, -1}, luma = -1;
    for (size_t c = 0; c < h.channels.size(); ++c) {
      const std::string& nm = h.channels[c].name;
      if (nm == "R") target[0] = (int)c;
      if (nm == "G") target[1] = (int)c;
      if (nm == "B") target[2] = (int)c;
      if (nm == "Y") luma = (int)c;
    }
    const bool rgb = target[0] >= 0 || target[1] >= 0 || target[2] >= 0;
    if (!rgb) {
      std::string names;
      for (const Channel& c : h.channels) names += (names.empty() ? "" : ", ") + c.name;
      supported(luma >= 0, "no R, G, B or Y channel (channels: " + names + ")");
    }
    std::vector<int> dest(h.channels.size(), -1);  // channel -> rgb slot, 3 = all three
    for (int t = 0; t < 3; ++t)
      if (target[t] >= 0) dest[target[t]] = t;
    if (!rgb) dest[luma] = 3;

    size_t line_bytes = 0;
    for (const Channel& c : h.channels) line_bytes += (size_t)width * c.bytes;
    out.rgb.assign((size_t)width * height * 3, 0.0f);
    std::vector<uint8_t> seen(chunks, 0), buf, tmp;
    for (uint32_t i = 0; i < chunks; ++i) {
      need(offsets[i] <= size && size - offsets[i] >= 8, "chunk offset out of the file");
      In cin(data, size, (size_t)offsets[i]);
      const int32_t y = cin.i32();
      const int32_t packed = cin.i32();
      need(y >= h.y0 && y <= h.y1 && ((int64_t)y - h.y0) % lpc == 0, "chunk has a bad first line");
      const uint32_t block = (uint32_t)(((int64_t)y - h.y0) / lpc);
      need(!seen[block], "two chunks hold the same lines");
      seen[block] = 1;
      need(packed >= 0 && (size_t)packed <= cin.left(), "chunk runs past the end of the file");
      const uint32_t lines = std::min<uint32_t>(lpc, (uint32_t)(h.y1 - y + 1));
      const size_t unpacked = (size_t)lines * line_bytes;
      const uint8_t* src = data + cin.pos;
      const uint8_t* px;
      if ((size_t)packed < unpacked) {
        buf.resize(unpacked);
        switch (h.compression) {
          case kRle:
            tmp.resize(unpacked);
            rle_unpack(src, (size_t)packed, tmp.data(), unpacked);
            unpredict_interleave(tmp.data(), buf.data(), unpacked);
            break;
          case kZips:
          case kZip:
            tmp.resize(unpacked);
            zip_unpack(src, (size_t)packed, tmp.data(), unpacked);
            unpredict_interleave(tmp.data(), buf.data(), unpacked);
            break;
          case kPiz:
            piz_unpack(src, (size_t)packed, width, lines, h.channels, buf.data(), unpacked);
            break;
          default:
            need(false, "uncompressed chunk with the wrong size");
        }
        px = buf.data();
      } else {
        need((size_t)packed == unpacked, "chunk larger than its unpacked size");
        px = src;
      }
      for (uint32_t l = 0; l < lines; ++l) {
        float* row = out.rgb.data() + ((size_t)(y - h.y0) + l) * width * 3;
        for (size_t c = 0; c < h.channels.size(); ++c) {
          const Channel& ch = h.channels[c];
          const int t = dest[c];
          if (t >= 0)
            for (uint32_t x = 0; x < width; ++x) {
              const uint8_t* s = px + (size_t)x * ch.bytes;
              float v;
              if (ch.type == kHalf) {
                v = half_to_float((uint16_t)(s[0] | (s[1] << 8)));
              } else if (ch.type == kFloat) {
                const uint32_t bits = rd32(s);
                std::memcpy(&v, &bits, 4);
              } else {
                v = (float)rd32(s);
              }
              if (t == 3)
                row[x * 3] = row[x * 3 + 1] = row[x * 3 + 2] = v;
              else
                row[x * 3 + t] = v;
            }
          px += (size_t)width * ch.bytes;
        }
      }
    }
    err.clear();
    return kExrOk;
  } catch (const Bad& b) {
    err = b.what();
    return b.kind;
  }
}

}  // namespace hpt
