// denoise_api.cpp — the learned denoiser behind the renderer's C ABI.
//
// The reference saves colour/albedo/normal PFMs and runs `oidnDenoise` on them
// (denoise.py:26-30, the Save button at src/bin/hala-pathtracer.rs:246-252).
// Here a host holding only include/hpt.h loads the trained U-Net once
// (hpt_denoiser_create over the bytes of the .npz file the training run
// writes) and denoises a renderer's resident images in place on the GPU
// (hpt_denoise), with the network's layer sequence driven from this file:
// five encoder levels of two 3x3 convolutions with 2x2 average pooling
// between them, four decoder levels whose first convolution reads the
// upsampled lower level concatenated with the skip (hpt_nn_conv_forward_upcat),
// a 1x1 output convolution plus the input's log colour, then
// max(expm1(.), 0).  The kernels are those of include/hpt_nn.h.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hpt.h"
#include "../../include/hpt_nn.h"

namespace hpt {
void set_last_error(const std::string& message);  // hpt_last_error() of this thread (api.cpp)
}

namespace {

constexpr int kInChannels = 16;
constexpr int kOutChannels = 3;

struct Fail : std::runtime_error {
  hpt_status code;
  Fail(hpt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void need(bool c, const std::string& m, hpt_status code = HPT_ERROR_INVALID_ARGUMENT) {
  if (!c) throw Fail(code, m);
}

void hip_ok(int e, const char* what) {
  if (e != 0) throw Fail(HPT_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString((hipError_t)e));
}

// ------------------------------------------------------------------ .npz
// A numpy .npz archive: a zip of .npy arrays (stored or deflated).  Only the
// central directory's entries are trusted for sizes and offsets.
uint32_t rd16(const uint8_t* p) { return p[0] | (p[1] << 8); }
uint32_t rd32(const uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

std::map<std::string, std::vector<uint8_t>> read_zip(const uint8_t* d, size_t n) {
  need(n >= 22, "weights: not a zip archive");
  size_t eocd = std::string::npos;
  for (size_t i = n - 22 + 1; i-- > 0 && n - i <= 22 + 65535;)
    if (rd32(d + i) == 0x06054b50u) {
      eocd = i;
      break;
    }
  need(eocd != std::string::npos, "weights: no zip end-of-directory record");
  const uint32_t count = rd16(d + eocd + 10);
  size_t cd = rd32(d + eocd + 16);
  std::map<std::string, std::vector<uint8_t>> out;
  for (uint32_t k = 0; k < count; ++k) {
    need(cd + 46 <= n && rd32(d + cd) == 0x02014b50u, "weights: bad zip directory entry");
    const uint32_t method = rd16(d + cd + 10);
    const size_t csize = rd32(d + cd + 20), usize = rd32(d + cd + 24);
    const size_t nlen = rd16(d + cd + 28), xlen = rd16(d + cd + 30), clen = rd16(d + cd + 32);
    const size_t local = rd32(d + cd + 42);
    need(csize != 0xFFFFFFFFu && usize != 0xFFFFFFFFu && local != 0xFFFFFFFFu, "weights: zip64 is not supported");
    need(cd + 46 + nlen <= n, "weights: truncated zip directory");
    const std::string name(reinterpret_cast<const char*>(d + cd + 46), nlen);
    need(local + 30 <= n && rd32(d + local) == 0x04034b50u, "weights: bad zip local header");
    const size_t data = local + 30 + rd16(d + local + 26) + rd16(d + local + 28);
    need(data + csize <= n, "weights: truncated zip entry " + name);
    std::vector<uint8_t> buf(usize);
    if (method == 0) {
      need(csize == usize, "weights: stored entry sizes differ");
      std::memcpy(buf.data(), d + data, usize);
    } else if (method == 8) {
      z_stream zs{};
      need(inflateInit2(&zs, -MAX_WBITS) == Z_OK, "weights: inflateInit2");
      zs.next_in = const_cast<uint8_t*>(d + data);
      zs.avail_in = (uInt)csize;
      zs.next_out = buf.data();
      zs.avail_out = (uInt)usize;
      const int rc = inflate(&zs, Z_FINISH);
      inflateEnd(&zs);
      need(rc == Z_STREAM_END && zs.total_out == usize, "weights: corrupt deflated entry " + name);
    } else {
      throw Fail(HPT_ERROR_UNSUPPORTED, "weights: zip compression method " + std::to_string(method));
    }
    out[name] = std::move(buf);
    cd += 46 + nlen + xlen + clen;
  }
  return out;
}

// One .npy array: little-endian dtype `descr` ("<f4", "<i8", ...), C order.
struct Npy {
  std::string descr;
  std::vector<size_t> shape;
  const uint8_t* data = nullptr;
  size_t bytes = 0;
};

Npy parse_npy(const std::vector<uint8_t>& b, const std::string& name) {
  need(b.size() >= 10 && std::memcmp(b.data(), "\x93NUMPY", 6) == 0, "weights: " + name + " is not a .npy array");
  const int major = b[6];
  size_t hlen, off;
  if (major == 1) {
    hlen = rd16(b.data() + 8);
    off = 10;
  } else {
    need(b.size() >= 12, "weights: truncated .npy header");
    hlen = rd32(b.data() + 8);
    off = 12;
  }
  need(off + hlen <= b.size(), "weights: truncated .npy header");
  const std::string h(reinterpret_cast<const char*>(b.data() + off), hlen);
  Npy a;
  auto field = [&](const char* key) -> size_t {
    const size_t p = h.find(key);
    need(p != std::string::npos, std::string("weights: .npy header without ") + key);
    return p + std::strlen(key);
  };
  size_t p = h.find('\'', field("'descr':"));
  a.descr = h.substr(p + 1, h.find('\'', p + 1) - p - 1);
  p = field("'fortran_order':");
  need(h.compare(h.find_first_not_of(' ', p), 5, "False") == 0, "weights: Fortran-order arrays are not supported");
  p = h.find('(', field("'shape':"));
  const size_t q = h.find(')', p);
  std::string dims = h.substr(p + 1, q - p - 1);
  size_t count = 1;
  for (size_t s = 0; s < dims.size();) {
    const size_t e = dims.find(',', s);
    const std::string tok = dims.substr(s, e == std::string::npos ? std::string::npos : e - s);
    if (tok.find_first_not_of(' ') != std::string::npos) {
      a.shape.push_back(std::stoull(tok));
      count *= a.shape.back();
    }
    if (e == std::string::npos) break;
    s = e + 1;
  }
  size_t item = 0;
  if (a.descr == "<f4" || a.descr == "<i4" || a.descr == "<u4") item = 4;
  else if (a.descr == "<i8" || a.descr == "<u8") item = 8;
  else throw Fail(HPT_ERROR_UNSUPPORTED, "weights: dtype " + a.descr + " of " + name);
  a.data = b.data() + off + hlen;
  a.bytes = count * item;
  need(off + hlen + a.bytes <= b.size(), "weights: truncated array " + name);
  return a;
}

int64_t npy_int(const Npy& a) {
  if (a.descr == "<i8" || a.descr == "<u8") {
    int64_t v;
    std::memcpy(&v, a.data, 8);
    return v;
  }
  int32_t v;
  std::memcpy(&v, a.data, 4);
  return v;
}

// ------------------------------------------------------------------ network
struct Layer {
  std::string name;
  int cin, cout, k, cout_pad;
  bool relu;
  size_t w_off, b_off;  // floats into the master parameter vector
  void* w = nullptr;    // bf16 [cout_pad][k*k][cin]
  void* wt = nullptr;   // bf16 [cin][k*k][cout_pad] (the data-gradient copy prepare_weights also writes)
};

int pad16(int c) { return (c + 15) / 16 * 16; }

}  // namespace

// The handle starts with a tag of its kind, so another kind of handle (or a
// destroyed one) is refused instead of written into.
constexpr uint32_t kDenoiserTag = 0x44545048u;  // "HPTD"

struct hpt_denoiser {
  uint32_t tag = kDenoiserTag;
  int32_t device = 0;
  hipStream_t stream = nullptr;
  std::vector<int> levels;
  bool residual = true;
  size_t param_count = 0;
  void* params = nullptr;  // float32 master parameters
  std::vector<Layer> layers;
  std::map<std::string, size_t> index;
  // Activation buffers of the last image size.
  int buf_h = 0, buf_w = 0;
  std::vector<void*> bufs;
  std::map<std::string, void*> buf;
  ~hpt_denoiser() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : bufs) (void)hipFree(p);
    for (Layer& L : layers) {
      if (L.w) (void)hipFree(L.w);
      if (L.wt) (void)hipFree(L.wt);
    }
    if (params) (void)hipFree(params);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <typename F>
hpt_status guarded(F&& fn) {
  try {
    fn();
    hpt::set_last_error("");
    return HPT_OK;
  } catch (const Fail& f) {
    hpt::set_last_error(f.what());
    return f.code;
  } catch (const std::bad_alloc&) {
    hpt::set_last_error("host allocation failed");
    return HPT_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    hpt::set_last_error(e.what());
    return HPT_ERROR_INVALID_ARGUMENT;
  }
}

void* dev_alloc(hpt_denoiser* d, size_t bytes) {
  void* p = nullptr;
  hip_ok(hpt_nn_alloc(bytes, &p), "hpt_nn_alloc");
  d->bufs.push_back(p);
  return p;
}

// Layers in forward order with their offsets in the parameter vector: the
// training code's layout (each layer's weights [cout][k*k][cin], then its
// cout biases).
void build_layers(hpt_denoiser* d) {
  const auto& lv = d->levels;
  const int L = (int)lv.size();
  auto add = [&](const std::string& name, int cin, int cout, int k, bool relu) {
    Layer l;
    l.name = name;
    l.cin = cin;
    l.cout = cout;
    l.k = k;
    l.relu = relu;
    l.cout_pad = pad16(cout);
    l.w_off = d->param_count;
    l.b_off = l.w_off + (size_t)cout * k * k * cin;
    d->param_count = l.b_off + cout;
    d->index[name] = d->layers.size();
    d->layers.push_back(l);
  };
  int cin = kInChannels;
  for (int l = 0; l < L; ++l) {
    add("enc" + std::to_string(l) + "a", cin, lv[l], 3, true);
    add("enc" + std::to_string(l) + "b", lv[l], lv[l], 3, true);
    cin = lv[l];
  }
  for (int l = L - 2; l >= 0; --l) {
    add("dec" + std::to_string(l) + "a", lv[l + 1] + lv[l], lv[l], 3, true);
    add("dec" + std::to_string(l) + "b", lv[l], lv[l], 3, true);
  }
  add("out", lv[0], kOutChannels, 1, false);
}

Layer& layer(hpt_denoiser* d, const std::string& name) { return d->layers[d->index.at(name)]; }

void ensure_buffers(hpt_denoiser* d, int h, int w) {
  if (d->buf_h == h && d->buf_w == w) return;
  hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  for (void* p : d->bufs) (void)hipFree(p);
  d->bufs.clear();
  d->buf.clear();
  d->buf_h = d->buf_w = 0;
  const int L = (int)d->levels.size();
  auto px = [&](int l) { return (size_t)(h >> l) * (size_t)(w >> l); };
  d->buf["x"] = dev_alloc(d, px(0) * kInChannels * 2);
  d->buf["out"] = dev_alloc(d, px(0) * kOutChannels * 4);
  for (const Layer& ly : d->layers) {
    if (ly.name == "out") continue;
    d->buf["y_" + ly.name] = dev_alloc(d, px(ly.name[3] - '0') * ly.cout * 2);
  }
  for (int l = 1; l < L; ++l) d->buf["pool" + std::to_string(l)] = dev_alloc(d, px(l) * d->levels[l - 1] * 2);
  d->buf_h = h;
  d->buf_w = w;
}

using u16 = uint16_t;
u16* B(hpt_denoiser* d, const std::string& k) { return static_cast<u16*>(d->buf.at(k)); }
const float* bias(hpt_denoiser* d, const Layer& l) { return static_cast<const float*>(d->params) + l.b_off; }

void conv(hpt_denoiser* d, const Layer& l, const u16* x, u16* y, float* y_f32, int h, int w) {
  hip_ok(hpt_nn_conv_forward(d->stream, x, static_cast<const u16*>(l.w), bias(d, l), y, y_f32, 1, h, w, l.cin, l.cout,
                             l.cout_pad, l.k, l.relu ? 1 : 0),
         ("convolution " + l.name).c_str());
}

// The inference forward pass on d->buf["x"] (1 x h x w x 16 bf16).
void forward(hpt_denoiser* d, int h, int w) {
  const int L = (int)d->levels.size();
  const u16* cur = B(d, "x");
  for (int l = 0; l < L; ++l) {
    const int hl = h >> l, wl = w >> l;
    const std::string s = std::to_string(l);
    if (l > 0) cur = B(d, "pool" + s);  // written by the previous level's second convolution
    conv(d, layer(d, "enc" + s + "a"), cur, B(d, "y_enc" + s + "a"), nullptr, hl, wl);
    const Layer& eb = layer(d, "enc" + s + "b");
    if (l + 1 < L)
      hip_ok(hpt_nn_conv_forward_pool(d->stream, B(d, "y_enc" + s + "a"), static_cast<const u16*>(eb.w), bias(d, eb),
                                      B(d, "y_enc" + s + "b"), B(d, "pool" + std::to_string(l + 1)), 1, hl, wl, eb.cin,
                                      eb.cout, eb.cout_pad, eb.relu ? 1 : 0),
             ("convolution " + eb.name).c_str());
    else
      conv(d, eb, B(d, "y_enc" + s + "a"), B(d, "y_enc" + s + "b"), nullptr, hl, wl);
    cur = B(d, "y_enc" + s + "b");
  }
  for (int l = L - 2; l >= 0; --l) {
    const int hl = h >> l, wl = w >> l;
    const std::string s = std::to_string(l);
    const Layer& a = layer(d, "dec" + s + "a");
    hip_ok(hpt_nn_conv_forward_upcat(d->stream, cur, B(d, "y_enc" + s + "b"), static_cast<const u16*>(a.w), bias(d, a),
                                     B(d, "y_dec" + s + "a"), nullptr, 1, hl, wl, d->levels[l + 1], d->levels[l],
                                     a.cout, a.cout_pad, a.relu ? 1 : 0),
           ("convolution " + a.name).c_str());
    conv(d, layer(d, "dec" + s + "b"), B(d, "y_dec" + s + "a"), B(d, "y_dec" + s + "b"), nullptr, hl, wl);
    cur = B(d, "y_dec" + s + "b");
  }
  float* out = static_cast<float*>(d->buf.at("out"));
  conv(d, layer(d, "out"), cur, nullptr, out, h, w);
  if (d->residual)
    hip_ok(hpt_nn_add_residual(d->stream, out, B(d, "x"), (size_t)h * w, kInChannels), "residual");
}

}  // namespace

extern "C" {

hpt_status hpt_denoiser_create(const void* npz, size_t size, int32_t device, hpt_denoiser** out) {
  hpt_denoiser* d = nullptr;
  const hpt_status st = guarded([&] {
    need(npz && out, "null argument");
    *out = nullptr;
    // The file is checked in full before any device is touched.
    const auto files = read_zip(static_cast<const uint8_t*>(npz), size);
    auto get = [&](const char* key) {
      const auto it = files.find(std::string(key) + ".npy");
      need(it != files.end(), std::string("weights: no array '") + key + "'");
      return parse_npy(it->second, key);
    };
    const Npy levels = get("levels"), params = get("params"), residual = get("residual");
    d = new hpt_denoiser;
    d->device = device;
    need(levels.shape.size() == 1 && levels.shape[0] >= 2 && levels.shape[0] <= 8, "weights: levels must hold 2-8 sizes");
    for (size_t i = 0; i < levels.shape[0]; ++i) {
      Npy one = levels;
      const size_t item = levels.bytes / levels.shape[0];
      one.data = levels.data + i * item;
      const int64_t c = npy_int(one);
      need(c > 0 && c <= 1024, "weights: level channel counts must be 1..1024");
      d->levels.push_back((int)c);
    }
    d->residual = npy_int(residual) != 0;
    build_layers(d);
    need(params.descr == "<f4", "weights: params must be float32");
    need(params.bytes == d->param_count * 4, "weights: " + std::to_string(params.bytes / 4) +
                                                 " parameters, the network has " + std::to_string(d->param_count));
    int n = 0;
    hip_ok(hipGetDeviceCount(&n), "hipGetDeviceCount");
    need(device >= 0 && device < n, "device ordinal out of range");
    hip_ok(hipSetDevice(device), "hipSetDevice");
    hip_ok(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_ok(hipMalloc(&d->params, params.bytes), "hipMalloc");
    hip_ok(hipMemcpyAsync(d->params, params.data, params.bytes, hipMemcpyHostToDevice, d->stream), "weights upload");
    for (Layer& l : d->layers) {
      hip_ok(hipMalloc(&l.w, (size_t)l.cout_pad * l.k * l.k * l.cin * 2), "hipMalloc");
      hip_ok(hipMalloc(&l.wt, (size_t)l.cin * l.k * l.k * l.cout_pad * 2), "hipMalloc");
      hip_ok(hpt_nn_prepare_weights(d->stream, static_cast<const float*>(d->params) + l.w_off, static_cast<u16*>(l.w),
                                    static_cast<u16*>(l.wt), l.cin, l.cout, l.cout_pad, l.k),
             "prepare weights");
    }
    hip_ok(hipStreamSynchronize(d->stream), "weights upload");
    *out = d;
  });
  if (st != HPT_OK) delete d;
  return st;
}

hpt_status hpt_denoise(hpt_denoiser* d, hpt_renderer* r, float* dst, size_t bytes, int dst_on_device) {
  return guarded([&] {
    need(d && d->tag == kDenoiserTag && r && dst, "null argument or invalid denoiser handle");
    hip_ok(hipSetDevice(d->device), "hipSetDevice");
    uint32_t width = 0, height = 0;
    if (hpt_get_image_size(r, &width, &height) != HPT_OK) throw Fail(HPT_ERROR_STATE, hpt_last_error());
    need(bytes == (size_t)width * height * 16, "destination size does not match the image size (width*height*16)");
    void* img[3];
    const hpt_image which[3] = {HPT_IMAGE_ACCUM_RGBA32F, HPT_IMAGE_ALBEDO_RGBA32F, HPT_IMAGE_NORMAL_RGBA32F};
    for (int i = 0; i < 3; ++i) {
      size_t nb = 0;
      if (hpt_get_image_pointer(r, which[i], &img[i], &nb) != HPT_OK) throw Fail(HPT_ERROR_STATE, hpt_last_error());
    }
    uint64_t spp = 0;
    if (hpt_get_frame_index(r, &spp) != HPT_OK || hpt_wait_idle(r) != HPT_OK)
      throw Fail(HPT_ERROR_STATE, hpt_last_error());
    const int step = 1 << ((int)d->levels.size() - 1);
    const int h = ((int)height + step - 1) / step * step, w = ((int)width + step - 1) / step * step;
    ensure_buffers(d, h, w);
    // the sample-count channel: log2(spp) / 8
    const float aux = (float)(std::log2((double)(spp > 0 ? spp : 1)) / 8.0);
    hip_ok(hpt_nn_make_input(d->stream, static_cast<const float*>(img[0]), static_cast<const float*>(img[1]),
                             static_cast<const float*>(img[2]), (int)width, (int)height, 0, 0, B(d, "x"), h, w, 0, aux),
           "network input");
    forward(d, h, w);
    // The denoised colour as float4 pixels (alpha 1): into the caller's
    // device image, or through a device buffer of the handle into host memory.
    // That buffer holds the padded h*w pixels, so every image size that
    // shares these activation buffers fits it.
    float* target = dst;
    if (!dst_on_device) {
      auto it = d->buf.find("result");
      target = static_cast<float*>(it != d->buf.end() ? it->second
                                                       : (d->buf["result"] = dev_alloc(d, (size_t)h * w * 16)));
    }
    hip_ok(hpt_nn_write_output(d->stream, static_cast<const float*>(d->buf.at("out")), target, (int)width,
                               (int)height, 0, 0, h, w),
           "output");
    if (!dst_on_device) hip_ok(hipMemcpyAsync(dst, target, bytes, hipMemcpyDeviceToHost, d->stream), "output copy");
    hip_ok(hipStreamSynchronize(d->stream), "denoise");
  });
}

void hpt_denoiser_destroy(hpt_denoiser* d) {
  if (!d || d->tag != kDenoiserTag) return;
  d->tag = 0;
  delete d;
}

}  // extern "C"
