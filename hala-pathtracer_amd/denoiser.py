"""Learned denoiser: a 5-level U-Net over the renderer's colour, albedo and
normal images, trained and run on the GPU by the kernels of include/hpt_nn.h
(BF16 convolutions on MFMA with float32 accumulation, hand-written forward
and backward).  It takes the place of the reference's offline OIDN step
(denoise.py:7-30, `oidnDenoise --hdr color.pfm --alb albedo.pfm --nrm
normal.pfm`): same three inputs, a denoised colour image out.

Network.  Input: log(1 + colour), albedo, normal * 0.5 + 0.5 and the sample
count's feature log2(spp) / 8 (10 channels, zero-padded to 16).  Encoder: two
3x3 convs + ReLU per level at 32, 48, 64, 96, 128 channels, 2x2 average
pooling between levels.  Decoder: nearest 2x upsampling, concatenation
[upsampled, skip], two 3x3 convs + ReLU per level.  Output: a 1x1 conv whose
result is added to the input's log(1 + colour) (the network predicts a
correction of its input).  Loss: L1 in that log space.

Precision.  Float32 master weights, gradients and Adam state; bf16 weights
and activations on the MFMA; float32 sums.  Images are NHWC on the device.

Memory.  Every buffer is device memory of libhpt.so's own HIP runtime
(DeviceArray over hpt_nn_alloc), and every operation is a call into the
library on one non-blocking stream of the process (stream()); there is no
CPU path.  Nothing runs on the legacy null stream, whose work would depend
on every blocking stream: a renderer capturing a HIP graph on another thread
(the server's jobs) must not see the denoiser's work join its capture.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

import hpt

LEVELS = (32, 48, 64, 96, 128)
IN_CHANNELS = 16
OUT_CHANNELS = 3
OUT_PAD = 16
_stream = None  # the denoiser's stream (non-blocking, like the renderers': wait for a renderer before reading its images)

_vp, _i, _sz, _f = C.c_void_p, C.c_int, C.c_size_t, C.c_float
# Entry points of include/hpt_nn.h with their argument types.
NN_SIGNATURES = {
    "hpt_nn_conv_forward": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i],
    "hpt_nn_conv_forward_upcat": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i],
    "hpt_nn_conv_forward_pool": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i],
    "hpt_nn_conv_backward_data": [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i],
    "hpt_nn_conv_backward_weights": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i],
    "hpt_nn_wgrad_scratch_floats": [_i, _i, _i, _i, _i, _i],
    "hpt_nn_prepare_weights": [_vp, _vp, _vp, _vp, _i, _i, _i, _i],
    "hpt_nn_relu_backward": [_vp, _vp, _vp, _sz],
    "hpt_nn_pool_forward": [_vp, _vp, _vp, _i, _i, _i, _i],
    "hpt_nn_pool_backward": [_vp, _vp, _vp, _i, _i, _i, _i],
    "hpt_nn_upsample_forward": [_vp, _vp, _vp, _i, _i, _i, _i],
    "hpt_nn_upsample_backward": [_vp, _vp, _vp, _i, _i, _i, _i],
    "hpt_nn_concat_forward": [_vp, _vp, _vp, _vp, _sz, _i, _i],
    "hpt_nn_concat_backward": [_vp, _vp, _vp, _vp, _sz, _i, _i],
    "hpt_nn_accumulate": [_vp, _vp, _vp, _sz],
    "hpt_nn_make_input": [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _i, _i, _i, _f],
    "hpt_nn_make_target": [_vp, _vp, _i, _i, _i, _i, _vp, _i, _i, _i],
    "hpt_nn_add_residual": [_vp, _vp, _vp, _sz, _i],
    "hpt_nn_l1_loss": [_vp, _vp, _vp, _vp, _vp, _sz, _i],
    "hpt_nn_adam": [_vp, _vp, _vp, _vp, _vp, _sz, _f, _f, _f, _f, _i],
    "hpt_nn_write_output": [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i],
    "hpt_nn_alloc": [_sz, C.POINTER(_vp)],
    "hpt_nn_free": [_vp],
    "hpt_nn_copy": [_vp, _vp, _vp, _sz],
    "hpt_nn_fill": [_vp, _vp, _f, _sz],
    "hpt_nn_stream_synchronize": [_vp],
    "hpt_nn_stream_create": [C.POINTER(_vp)],
    "hpt_nn_event_create": [C.POINTER(_vp)],
    "hpt_nn_event_destroy": [_vp],
    "hpt_nn_event_record": [_vp, _vp],
    "hpt_nn_event_elapsed": [_vp, _vp, C.POINTER(_f)],
}

_nn = None


def nn_library():
    """libhpt.so with the denoiser entry points typed; raises if not built."""
    global _nn
    if _nn is None:
        lib = hpt.load_library()
        for name, args in NN_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_size_t if name == "hpt_nn_wgrad_scratch_floats" else C.c_int
        _nn = lib
    return _nn


def _call(name, *args):
    rc = getattr(nn_library(), name)(*args)
    if rc:
        raise RuntimeError(f"{name} failed with HIP error {rc}")


def stream():
    """The process's denoiser stream, created on first use."""
    global _stream
    if _stream is None:
        p = C.c_void_p()
        _call("hpt_nn_stream_create", C.byref(p))
        _stream = p
    return _stream


def synchronize() -> None:
    """Waits for the denoiser's work (its stream)."""
    _call("hpt_nn_stream_synchronize", stream())


class DeviceArray:
    """A flat device buffer of `count` elements of a numpy dtype, allocated
    (zeroed) by libhpt.so and freed with the object."""

    def __init__(self, count: int, dtype, shape=None):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.shape = tuple(shape) if shape is not None else (self.count,)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        _call("hpt_nn_alloc", self.nbytes, C.byref(p))
        self.ptr = int(p.value or 0)

    def __del__(self):
        if getattr(self, "ptr", 0) and _nn is not None:
            _nn.hpt_nn_free(C.c_void_p(self.ptr))
            self.ptr = 0

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, a.shape)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        """Copies host data in at element `offset` (synchronously)."""
        a = np.ascontiguousarray(a, self.dtype)
        if offset + a.size > self.count:
            raise ValueError("upload past the end of the buffer")
        _call("hpt_nn_copy", stream(), C.c_void_p(self.addr(offset)), C.c_void_p(a.ctypes.data), a.nbytes)
        synchronize()

    def copy_from_device(self, src: int, nbytes: int, offset: int = 0) -> None:
        """Device-to-device copy from address `src` (queued on the stream)."""
        if offset * self.dtype.itemsize + nbytes > self.nbytes:
            raise ValueError("copy past the end of the buffer")
        _call("hpt_nn_copy", stream(), C.c_void_p(self.addr(offset)), C.c_void_p(src), nbytes)

    def numpy(self) -> np.ndarray:
        """A host copy, after the queued work has finished."""
        out = np.empty(self.count, self.dtype)
        _call("hpt_nn_copy", stream(), C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), self.nbytes)
        synchronize()
        return out.reshape(self.shape)

    def fill(self, value: float) -> None:
        if self.dtype != np.float32:
            raise TypeError("fill is for float32 buffers")
        _call("hpt_nn_fill", stream(), C.c_void_p(self.ptr), value, self.count)

    def addr(self, offset: int = 0) -> int:
        """Device address of element `offset`."""
        if not 0 <= offset <= self.count:
            raise IndexError(offset)
        return self.ptr + offset * self.dtype.itemsize


class _Ptr(C.c_void_p):
    """A pointer argument that holds its DeviceArray until the call returns
    (a temporary freed while the arguments are evaluated would hand its
    memory to the next argument's allocation)."""


def _p(t):
    """A device pointer argument: a DeviceArray, a raw address (int) or None."""
    if t is None:
        return None
    if isinstance(t, DeviceArray):
        v = _Ptr(t.ptr)
        v.keep = t
        return v
    return C.c_void_p(int(t))


class Event:
    """A HIP event of the library's runtime."""

    def __init__(self):
        e = C.c_void_p()
        _call("hpt_nn_event_create", C.byref(e))
        self.e = e

    def __del__(self):
        if getattr(self, "e", None) and self.e.value and _nn is not None:
            _nn.hpt_nn_event_destroy(self.e)

    def record(self, on=None) -> None:
        """Records the event on stream `on` (default: the denoiser's)."""
        _call("hpt_nn_event_record", self.e, on if on is not None else stream())

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        _call("hpt_nn_event_elapsed", self.e, stop.e, C.byref(ms))
        return float(ms.value)


def pad16(c: int) -> int:
    return (c + 15) // 16 * 16


def spp_feature(spp: int) -> float:
    """The sample-count input channel: log2(spp) / 8."""
    return math.log2(max(1, spp)) / 8.0


def layer_specs(levels=LEVELS):
    """[(name, cin, cout, k, relu)] in forward order (see the module doc)."""
    specs = []
    cin = IN_CHANNELS
    for l, c in enumerate(levels):
        specs += [(f"enc{l}a", cin, c, 3, True), (f"enc{l}b", c, c, 3, True)]
        cin = c
    for l in range(len(levels) - 2, -1, -1):
        specs += [(f"dec{l}a", levels[l + 1] + levels[l], levels[l], 3, True),
                  (f"dec{l}b", levels[l], levels[l], 3, True)]
    specs.append(("out", levels[0], OUT_CHANNELS, 1, False))
    return specs


class _Layer:
    def __init__(self, name, cin, cout, k, relu, offset):
        self.name, self.cin, self.cout, self.k, self.relu = name, cin, cout, k, relu
        self.cout_pad = pad16(cout)
        self.w_size = cout * k * k * cin
        self.w_off, self.b_off = offset, offset + self.w_size
        self.size = self.w_size + cout
        self.w = DeviceArray(self.cout_pad * k * k * cin, np.uint16)
        self.wt = DeviceArray(cin * k * k * self.cout_pad, np.uint16)


class UNet:
    """The denoising network with its optimizer state, on the current GPU."""

    def __init__(self, levels=LEVELS, seed: int = 0, residual: bool = True):
        nn_library()
        self.levels = tuple(levels)
        self.residual = residual
        self.layers = {}
        off = 0
        for name, cin, cout, k, relu in layer_specs(self.levels):
            self.layers[name] = _Layer(name, cin, cout, k, relu, off)
            off += self.layers[name].size
        self.param_count = off
        self.params = DeviceArray(off, np.float32)
        self.grads = DeviceArray(off, np.float32)
        self.adam_m = DeviceArray(off, np.float32)
        self.adam_v = DeviceArray(off, np.float32)
        self.step_count = 0
        self._bufs = {}
        self._scratch = None
        self._loss = DeviceArray(1, np.float32)
        self.set_params(self.init_params(seed))

    # -------------------------------------------------------- parameters
    def init_params(self, seed: int) -> dict:
        """He-normal weights [cout][k*k][cin], biases 0.01 (numpy, float32)."""
        rng = np.random.default_rng(seed)
        out = {}
        for name, cin, cout, k, _ in layer_specs(self.levels):
            std = math.sqrt(2.0 / (k * k * cin))
            out[name] = (rng.normal(0.0, std, (cout, k * k, cin)).astype(np.float32), np.full(cout, 0.01, np.float32))
        return out

    def set_params(self, params: dict) -> None:
        flat = np.zeros(self.param_count, np.float32)
        for name, L in self.layers.items():
            w, b = params[name]
            flat[L.w_off:L.b_off] = np.asarray(w, np.float32).reshape(-1)
            flat[L.b_off:L.b_off + L.cout] = np.asarray(b, np.float32).reshape(-1)
        self.params.upload(flat)
        self.prepare_weights()

    def _unflatten(self, flat: np.ndarray) -> dict:
        return {name: (flat[L.w_off:L.b_off].reshape(L.cout, L.k * L.k, L.cin).copy(),
                       flat[L.b_off:L.b_off + L.cout].copy()) for name, L in self.layers.items()}

    def get_params(self) -> dict:
        return self._unflatten(self.params.numpy())

    def get_grads(self) -> dict:
        return self._unflatten(self.grads.numpy())

    def prepare_weights(self) -> None:
        """Master float32 weights -> the bf16 copies the convolutions read."""
        for L in self.layers.values():
            _call("hpt_nn_prepare_weights", stream(), _p(self.params.addr(L.w_off)), _p(L.w), _p(L.wt), L.cin, L.cout,
                  L.cout_pad, L.k)

    def save(self, path: str) -> None:
        np.savez(path, levels=np.array(self.levels), params=self.params.numpy(),
                 step=np.array(self.step_count), residual=np.array(int(self.residual)))

    @classmethod
    def load(cls, path: str) -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            net = cls(tuple(int(c) for c in f["levels"]), residual=bool(int(f["residual"])))
            params = f["params"]
            if params.size != net.param_count:
                raise ValueError(f"{path}: {params.size} parameters, the network has {net.param_count}")
            net.params.upload(params.astype(np.float32))
            net.step_count = int(f["step"])
        net.prepare_weights()
        return net

    # -------------------------------------------------------- buffers
    def _buffers(self, n: int, h: int, w: int, train: bool) -> dict:
        key = (n, h, w, train)
        if key in self._bufs:
            return self._bufs[key]
        step = 1 << (len(self.levels) - 1)
        if h % step or w % step:
            raise ValueError(f"image size {w}x{h} is not a multiple of {step}")
        bf = lambda px, c: DeviceArray(px * c, np.uint16)  # noqa: E731
        f32 = lambda px, c: DeviceArray(px * c, np.float32)  # noqa: E731
        P = [n * (h >> l) * (w >> l) for l in range(len(self.levels))]
        b = {"x": bf(P[0], IN_CHANNELS), "out": f32(P[0], OUT_CHANNELS)}
        for name, L in self.layers.items():
            if name != "out":
                b["y_" + name] = bf(P[self._level(name)], L.cout)
        for l in range(1, len(self.levels)):
            b[f"pool{l}"] = bf(P[l], self.levels[l - 1])
        for l in range(len(self.levels) - 1):
            b[f"up{l}"] = bf(P[l], self.levels[l + 1])
            b[f"cat{l}"] = bf(P[l], self.levels[l + 1] + self.levels[l])
        if train:
            b["target"] = f32(P[0], OUT_CHANNELS)
            b["dout"] = f32(P[0], OUT_PAD)
            for name, L in self.layers.items():
                b["dx_" + name] = f32(P[self._level(name)], L.cin)
            for l in range(len(self.levels)):
                b[f"dlev{l}"] = f32(P[l], self.levels[l])  # gradient w.r.t. a level's output
                b[f"dpool{l}"] = f32(P[l], self.levels[l])
            for l in range(len(self.levels) - 1):
                b[f"dup{l}"] = f32(P[l], self.levels[l + 1])
            need = max(nn_library().hpt_nn_wgrad_scratch_floats(n, h >> self._level(nm), w >> self._level(nm), L.cin,
                                                                  L.cout_pad, L.k)
                       for nm, L in self.layers.items())
            if self._scratch is None or self._scratch.count < need:
                self._scratch = DeviceArray(need, np.float32)
        b["P"] = P
        self._bufs[key] = b
        return b

    @staticmethod
    def _level(name: str) -> int:
        return 0 if name == "out" else int(name[3])

    # -------------------------------------------------------- passes
    def _conv(self, L, x, y, y_f32, n, h, w):
        _call("hpt_nn_conv_forward", stream(), _p(x), _p(L.w), _p(self.params.addr(L.b_off)), _p(y), _p(y_f32), n, h, w,
              L.cin, L.cout, L.cout_pad, L.k, int(L.relu))

    def _conv_pool(self, L, x, y, y_pool, n, h, w):
        """L (3x3) into y and its 2x2 average pooling into y_pool."""
        _call("hpt_nn_conv_forward_pool", stream(), _p(x), _p(L.w), _p(self.params.addr(L.b_off)), _p(y), _p(y_pool),
              n, h, w, L.cin, L.cout, L.cout_pad, int(L.relu))

    def _conv_upcat(self, L, low, skip, y, n, h, w, c_low, c_skip):
        """L over concat(upsample2x(low), skip) without materialising it."""
        _call("hpt_nn_conv_forward_upcat", stream(), _p(low), _p(skip), _p(L.w), _p(self.params.addr(L.b_off)), _p(y),
              None, n, h, w, c_low, c_skip, L.cout, L.cout_pad, int(L.relu))

    def forward_input(self, x: DeviceArray, n: int, h: int, w: int, train: bool = False) -> DeviceArray:
        """Runs the network on the bf16 input x (n*h*w*16 bit patterns, NHWC;
        copied into the network's input buffer unless it is that buffer);
        returns the float32 output buffer (n*h*w*3)."""
        b = self._buffers(n, h, w, train)
        if x.ptr != b["x"].ptr:
            b["x"].copy_from_device(x.ptr, b["x"].nbytes)
        st = stream()
        Lv = len(self.levels)
        cur = b["x"]
        for l in range(Lv):
            hl, wl = h >> l, w >> l
            if l > 0:
                cur = b[f"pool{l}"]  # written by the previous level's second convolution
            self._conv(self.layers[f"enc{l}a"], cur, b[f"y_enc{l}a"], None, n, hl, wl)
            if l + 1 < Lv:
                self._conv_pool(self.layers[f"enc{l}b"], b[f"y_enc{l}a"], b[f"y_enc{l}b"], b[f"pool{l + 1}"], n, hl, wl)
            else:
                self._conv(self.layers[f"enc{l}b"], b[f"y_enc{l}a"], b[f"y_enc{l}b"], None, n, hl, wl)
            cur = b[f"y_enc{l}b"]
        for l in range(Lv - 2, -1, -1):
            hl, wl = h >> l, w >> l
            if train:  # the backward pass reads the concatenated input
                _call("hpt_nn_upsample_forward", st, _p(cur), _p(b[f"up{l}"]), n, hl // 2, wl // 2, self.levels[l + 1])
                _call("hpt_nn_concat_forward", st, _p(b[f"up{l}"]), _p(b[f"y_enc{l}b"]), _p(b[f"cat{l}"]),
                      n * hl * wl, self.levels[l + 1], self.levels[l])
                self._conv(self.layers[f"dec{l}a"], b[f"cat{l}"], b[f"y_dec{l}a"], None, n, hl, wl)
            else:
                self._conv_upcat(self.layers[f"dec{l}a"], cur, b[f"y_enc{l}b"], b[f"y_dec{l}a"], n, hl, wl,
                                 self.levels[l + 1], self.levels[l])
            self._conv(self.layers[f"dec{l}b"], b[f"y_dec{l}a"], b[f"y_dec{l}b"], None, n, hl, wl)
            cur = b[f"y_dec{l}b"]
        self._conv(self.layers["out"], cur, None, b["out"], n, h, w)
        if self.residual:
            _call("hpt_nn_add_residual", st, _p(b["out"]), _p(b["x"]), n * h * w, IN_CHANNELS)
        return b["out"]

    def _conv_back(self, L, dy, x, y, dx, n, h, w):
        """dy: gradient w.r.t. the layer's output (modified in place by the
        ReLU's derivative); weight and bias gradients into self.grads; the
        input gradient into dx when dx is given."""
        st = stream()
        if L.relu:
            _call("hpt_nn_relu_backward", st, _p(dy), _p(y), n * h * w * L.cout)
        _call("hpt_nn_conv_backward_weights", st, _p(dy), _p(x), _p(self.grads.addr(L.w_off)),
              _p(self.grads.addr(L.b_off)), _p(self._scratch), n, h, w, L.cin, L.cout, L.cout_pad, L.k)
        if dx is not None:
            _call("hpt_nn_conv_backward_data", st, _p(dy), _p(L.wt), _p(dx), n, h, w, L.cin, L.cout, L.cout_pad, L.k)

    def backward(self, n: int, h: int, w: int) -> None:
        """Gradients of the loss whose output gradient is in the "dout"
        buffer (float32, 16 channels per pixel) into self.grads."""
        b = self._buffers(n, h, w, True)
        st = stream()
        Lv = len(self.levels)
        lay = self.layers
        self._conv_back(lay["out"], b["dout"], b["y_dec0b"] if Lv > 1 else b["y_enc0b"], None, b["dx_out"], n, h, w)
        d = b["dx_out"]
        for l in range(Lv - 1):
            hl, wl = h >> l, w >> l
            self._conv_back(lay[f"dec{l}b"], d, b[f"y_dec{l}a"], b[f"y_dec{l}b"], b[f"dx_dec{l}b"], n, hl, wl)
            self._conv_back(lay[f"dec{l}a"], b[f"dx_dec{l}b"], b[f"cat{l}"], b[f"y_dec{l}a"], b[f"dx_dec{l}a"], n,
                            hl, wl)
            _call("hpt_nn_concat_backward", st, _p(b[f"dx_dec{l}a"]), _p(b[f"dup{l}"]), _p(b[f"dlev{l}"]),
                  n * hl * wl, self.levels[l + 1], self.levels[l])
            _call("hpt_nn_upsample_backward", st, _p(b[f"dup{l}"]), _p(b[f"dlev{l + 1}"]), n, hl // 2, wl // 2,
                  self.levels[l + 1])
            d = b[f"dlev{l + 1}"]
        for l in range(Lv - 1, -1, -1):
            hl, wl = h >> l, w >> l
            if l < Lv - 1:
                _call("hpt_nn_pool_backward", st, _p(b[f"dx_enc{l + 1}a"]), _p(b[f"dpool{l}"]), n, hl, wl,
                      self.levels[l])
                _call("hpt_nn_accumulate", st, _p(b[f"dlev{l}"]), _p(b[f"dpool{l}"]), n * hl * wl * self.levels[l])
            x_in = b["x"] if l == 0 else b[f"pool{l}"]
            self._conv_back(lay[f"enc{l}b"], b[f"dlev{l}"], b[f"y_enc{l}a"], b[f"y_enc{l}b"], b[f"dx_enc{l}b"], n,
                            hl, wl)
            self._conv_back(lay[f"enc{l}a"], b[f"dx_enc{l}b"], x_in, b[f"y_enc{l}a"],
                            b[f"dx_enc{l}a"] if l > 0 else None, n, hl, wl)

    def loss(self, out: DeviceArray, target: DeviceArray, n: int, h: int, w: int, loss_sum: DeviceArray) -> None:
        """L1 loss of `out` against `target` (log colours): its mean is
        added to loss_sum[0] and its gradient written to the "dout" buffer."""
        b = self._buffers(n, h, w, True)
        _call("hpt_nn_l1_loss", stream(), _p(out), _p(target), _p(b["dout"]), _p(loss_sum), n * h * w, OUT_PAD)

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_count += 1
        _call("hpt_nn_adam", stream(), _p(self.params), _p(self.grads), _p(self.adam_m), _p(self.adam_v),
              self.param_count, lr, beta1, beta2, eps, self.step_count)
        self.prepare_weights()

    def train_step(self, x: DeviceArray, target: DeviceArray, n: int, h: int, w: int, lr: float,
                   loss_sum: DeviceArray | None = None) -> float | None:
        """One step on a batch: forward, L1 loss in log space, backward, Adam.
        The loss is added to loss_sum[0] when given (nothing waits); else it
        is returned (the step is waited for)."""
        b = self._buffers(n, h, w, True)
        out = self.forward_input(x, n, h, w, train=True)
        if target.ptr != b["target"].ptr:
            b["target"].copy_from_device(target.ptr, b["target"].nbytes)
        acc = loss_sum
        if acc is None:
            acc = self._loss
            acc.fill(0.0)
        self.loss(out, b["target"], n, h, w, acc)
        self.backward(n, h, w)
        self.adam_step(lr)
        return None if loss_sum is not None else float(acc.numpy()[0])

    # -------------------------------------------------------- inference
    def make_input(self, images, src_w: int, src_h: int, x0: int, y0: int, h: int, w: int, out, spp: int,
                   mirror: int = 0):
        """Network input (h*w*16 bf16) into `out` (a buffer or an address)
        of the h x w window at (x0, y0) of (colour, albedo, normal): float4
        device images of src_w x src_h averaged over `spp` samples, given as
        DeviceArrays or addresses."""
        color, albedo, normal = images
        _call("hpt_nn_make_input", stream(), _p(color), _p(albedo), _p(normal), src_w, src_h, x0, y0, _p(out), h, w,
              mirror, spp_feature(spp))
        return out

    def make_target(self, color, src_w: int, src_h: int, x0: int, y0: int, h: int, w: int, out, mirror: int = 0):
        """log(1 + colour) of the same window of a reference image into `out`
        (h*w*3 float32)."""
        _call("hpt_nn_make_target", stream(), _p(color), src_w, src_h, x0, y0, _p(out), h, w, mirror)
        return out

    def denoise(self, images, width: int, height: int, spp: int, out: DeviceArray | None = None) -> DeviceArray:
        """Denoised colour (height, width, 4) float32 of (colour, albedo,
        normal) float4 device images (DeviceArrays or addresses) of `spp`
        samples per pixel; the network runs on the image padded up to a
        multiple of 16 (zero features outside)."""
        step = 1 << (len(self.levels) - 1)
        h, w = -(-height // step) * step, -(-width // step) * step
        b = self._buffers(1, h, w, False)
        self.make_input(images, width, height, 0, 0, h, w, b["x"], spp)
        res = self.forward_input(b["x"], 1, h, w)
        if out is None:
            out = DeviceArray(height * width * 4, np.float32, (height, width, 4))
        _call("hpt_nn_write_output", stream(), _p(res), _p(out), width, height, 0, 0, h, w)
        return out

    def denoise_renderer(self, r: "hpt.HalaRenderer", out: DeviceArray | None = None) -> DeviceArray:
        """Denoises a renderer's accumulated images in place on the device
        (the role of denoise.py's oidnDenoise call on the saved PFMs)."""
        r.wait_idle()
        images = tuple(r.image_pointer(k)[0] for k in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL))
        return self.denoise(images, r.width, r.height, max(1, r.frame_index), out)

    def profile_forward(self, n: int, h: int, w: int, reps: int = 5) -> dict:
        """{layer: (average ms, algorithmic FLOP)} of each convolution of the
        forward pass at n x h x w (an encoder's second convolution with the
        pooling it writes), timed with HIP events on the stream the kernels
        run on; the input buffer is used as it is."""
        b = self._buffers(n, h, w, False)
        ev = {name: [] for name in self.layers}

        def timed(orig):
            def run(L, *args):
                s, e = Event(), Event()
                s.record()
                orig(L, *args)
                e.record()
                ev[L.name].append((s, e))
            return run

        self._conv, self._conv_upcat = timed(self._conv), timed(self._conv_upcat)
        self._conv_pool = timed(self._conv_pool)
        try:
            for _ in range(reps):
                self.forward_input(b["x"], n, h, w)
        finally:
            del self._conv, self._conv_upcat, self._conv_pool
        synchronize()
        out = {}
        for name, L in self.layers.items():
            l = self._level(name)
            flops = 2 * n * (h >> l) * (w >> l) * L.k * L.k * L.cin * L.cout
            out[name] = (sum(s.elapsed_ms(e) for s, e in ev[name]) / len(ev[name]), flops)
        return out

    def flops_forward(self, n: int, h: int, w: int) -> int:
        """Multiply-adds x 2 of the convolutions in one forward pass."""
        total = 0
        for name, L in self.layers.items():
            l = self._level(name)
            total += 2 * n * (h >> l) * (w >> l) * L.k * L.k * L.cin * L.cout
        return total


def renderer_images(r: "hpt.HalaRenderer") -> tuple:
    """Copies of a renderer's accumulated (colour, albedo, normal) images
    (after wait_idle), device to device."""
    r.wait_idle()
    out = []
    for k in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL):
        ptr, nbytes = r.image_pointer(k)
        d = DeviceArray(r.height * r.width * 4, np.float32, (r.height, r.width, 4))
        d.copy_from_device(ptr, nbytes)
        out.append(d)
    synchronize()  # the renderer may overwrite its images once it renders again
    return tuple(out)


def weights_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_denoiser.npz")
