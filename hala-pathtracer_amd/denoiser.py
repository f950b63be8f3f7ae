"""Learned denoiser: a 5-level U-Net over the renderer's colour, albedo and
normal images, trained and run on the GPU by the kernels of include/hpt_nn.h
(BF16 convolutions on MFMA with float32 accumulation, hand-written forward
and backward).  It takes the place of the reference's offline OIDN step
(denoise.py:7-30, `oidnDenoise --hdr color.pfm --alb albedo.pfm --nrm
normal.pfm`): same three inputs, a denoised colour image out.

Network.  Input: log(1 + colour), albedo, normal * 0.5 + 0.5 (9 channels,
zero-padded to 16).  Encoder: two 3x3 convs + ReLU per level at 32, 48, 64,
96, 128 channels, 2x2 average pooling between levels.  Decoder: nearest 2x
upsampling, concatenation [upsampled, skip], two 3x3 convs + ReLU per level.
Output: a 1x1 conv to log(1 + colour).  Loss: L1 in that log space.

Precision.  Float32 master weights, gradients and Adam state; bf16 weights
and activations on the MFMA; float32 sums.  Images are NHWC on the device.

torch provides device memory and the current stream only; every operation is
a call into libhpt.so.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import torch

import hpt

LEVELS = (32, 48, 64, 96, 128)
IN_CHANNELS = 16
OUT_CHANNELS = 3
OUT_PAD = 16

_vp, _i, _sz, _f = C.c_void_p, C.c_int, C.c_size_t, C.c_float
# Entry points of include/hpt_nn.h with their argument types.
NN_SIGNATURES = {
    "hpt_nn_conv_forward": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i],
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
    "hpt_nn_make_input": [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _i, _i],
    "hpt_nn_l1_loss": [_vp, _vp, _vp, _vp, _vp, _sz, _i],
    "hpt_nn_adam": [_vp, _vp, _vp, _vp, _vp, _sz, _f, _f, _f, _f, _i],
    "hpt_nn_write_output": [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i],
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


class _Ptr(C.c_void_p):
    """A device pointer holding its tensor until the call returns (a tensor
    freed while the arguments are evaluated could be reused by the next
    argument's allocation before the kernel reads it)."""


def _p(t):
    if t is None:
        return None
    v = _Ptr(t.data_ptr())
    v.keep = t
    return v


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def pad16(c: int) -> int:
    return (c + 15) // 16 * 16


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
    def __init__(self, name, cin, cout, k, relu, offset, device):
        self.name, self.cin, self.cout, self.k, self.relu = name, cin, cout, k, relu
        self.cout_pad = pad16(cout)
        self.w_size = cout * k * k * cin
        self.w_off, self.b_off = offset, offset + self.w_size
        self.size = self.w_size + cout
        self.w = torch.empty(self.cout_pad * k * k * cin, dtype=torch.int16, device=device)
        self.wt = torch.empty(cin * k * k * self.cout_pad, dtype=torch.int16, device=device)


class UNet:
    """The denoising network with its optimizer state, on one GPU."""

    def __init__(self, levels=LEVELS, seed: int = 0, device: int = 0):
        nn_library()
        self.levels = tuple(levels)
        self.device = torch.device("cuda", device)
        self.layers = {}
        off = 0
        for name, cin, cout, k, relu in layer_specs(self.levels):
            self.layers[name] = _Layer(name, cin, cout, k, relu, off, self.device)
            off += self.layers[name].size
        self.param_count = off
        self.params = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.step_count = 0
        self._bufs = {}
        self._scratch = torch.empty(0, dtype=torch.float32, device=self.device)
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)
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
        self.params.copy_(torch.from_numpy(flat))
        self.prepare_weights()

    def get_params(self) -> dict:
        flat = self.params.cpu().numpy()
        return {name: (flat[L.w_off:L.b_off].reshape(L.cout, L.k * L.k, L.cin).copy(),
                       flat[L.b_off:L.b_off + L.cout].copy()) for name, L in self.layers.items()}

    def get_grads(self) -> dict:
        flat = self.grads.cpu().numpy()
        return {name: (flat[L.w_off:L.b_off].reshape(L.cout, L.k * L.k, L.cin).copy(),
                       flat[L.b_off:L.b_off + L.cout].copy()) for name, L in self.layers.items()}

    def prepare_weights(self) -> None:
        """Master float32 weights -> the bf16 copies the convolutions read."""
        st = _stream()
        for L in self.layers.values():
            _call("hpt_nn_prepare_weights", st, _p(self.params[L.w_off:]), _p(L.w), _p(L.wt), L.cin, L.cout,
                  L.cout_pad, L.k)

    def save(self, path: str) -> None:
        np.savez(path, levels=np.array(self.levels), params=self.params.cpu().numpy(),
                 step=np.array(self.step_count))

    @classmethod
    def load(cls, path: str, device: int = 0) -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            net = cls(tuple(int(c) for c in f["levels"]), device=device)
            params = f["params"]
            if params.size != net.param_count:
                raise ValueError(f"{path}: {params.size} parameters, the network has {net.param_count}")
            net.params.copy_(torch.from_numpy(params.astype(np.float32)))
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
        dev = self.device
        bf = lambda px, c: torch.empty(px * c, dtype=torch.int16, device=dev)  # noqa: E731
        f32 = lambda px, c: torch.empty(px * c, dtype=torch.float32, device=dev)  # noqa: E731
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
            if self._scratch.numel() < need:
                self._scratch = torch.empty(need, dtype=torch.float32, device=dev)
        b["P"] = P
        self._bufs[key] = b
        return b

    @staticmethod
    def _level(name: str) -> int:
        return 0 if name == "out" else int(name[3])

    # -------------------------------------------------------- passes
    def _conv(self, L, x, y, y_f32, n, h, w):
        _call("hpt_nn_conv_forward", _stream(), _p(x), _p(L.w), _p(self.params[L.b_off:]), _p(y), _p(y_f32), n, h, w,
              L.cin, L.cout, L.cout_pad, L.k, int(L.relu))

    def forward_input(self, x: torch.Tensor, n: int, h: int, w: int, train: bool = False) -> torch.Tensor:
        """Runs the network on the bf16 input x ((n*h*w*16) int16 bit
        patterns, NHWC); returns the float32 output (n*h*w*3)."""
        b = self._buffers(n, h, w, train)
        if x.data_ptr() != b["x"].data_ptr():
            b["x"].copy_(x.reshape(-1))
        st = _stream()
        Lv = len(self.levels)
        cur = b["x"]
        for l in range(Lv):
            hl, wl = h >> l, w >> l
            if l > 0:
                _call("hpt_nn_pool_forward", st, _p(b[f"y_enc{l - 1}b"]), _p(b[f"pool{l}"]), n, 2 * hl, 2 * wl,
                      self.levels[l - 1])
                cur = b[f"pool{l}"]
            self._conv(self.layers[f"enc{l}a"], cur, b[f"y_enc{l}a"], None, n, hl, wl)
            self._conv(self.layers[f"enc{l}b"], b[f"y_enc{l}a"], b[f"y_enc{l}b"], None, n, hl, wl)
            cur = b[f"y_enc{l}b"]
        for l in range(Lv - 2, -1, -1):
            hl, wl = h >> l, w >> l
            _call("hpt_nn_upsample_forward", st, _p(cur), _p(b[f"up{l}"]), n, hl // 2, wl // 2, self.levels[l + 1])
            _call("hpt_nn_concat_forward", st, _p(b[f"up{l}"]), _p(b[f"y_enc{l}b"]), _p(b[f"cat{l}"]),
                  n * hl * wl, self.levels[l + 1], self.levels[l])
            self._conv(self.layers[f"dec{l}a"], b[f"cat{l}"], b[f"y_dec{l}a"], None, n, hl, wl)
            self._conv(self.layers[f"dec{l}b"], b[f"y_dec{l}a"], b[f"y_dec{l}b"], None, n, hl, wl)
            cur = b[f"y_dec{l}b"]
        self._conv(self.layers["out"], cur, None, b["out"], n, h, w)
        return b["out"]

    def _conv_back(self, L, dy, x, y, dx, n, h, w):
        """dy: gradient w.r.t. the layer's output (modified in place by the
        ReLU's derivative); weight and bias gradients into self.grads; the
        input gradient into dx when dx is given."""
        st = _stream()
        if L.relu:
            _call("hpt_nn_relu_backward", st, _p(dy), _p(y), n * h * w * L.cout)
        _call("hpt_nn_conv_backward_weights", st, _p(dy), _p(x), _p(self.grads[L.w_off:]), _p(self.grads[L.b_off:]),
              _p(self._scratch), n, h, w, L.cin, L.cout, L.cout_pad, L.k)
        if dx is not None:
            _call("hpt_nn_conv_backward_data", st, _p(dy), _p(L.wt), _p(dx), n, h, w, L.cin, L.cout, L.cout_pad, L.k)

    def backward(self, n: int, h: int, w: int) -> None:
        """Gradients of the loss whose output gradient is in the "dout"
        buffer (float32, 16 channels per pixel) into self.grads."""
        b = self._buffers(n, h, w, True)
        st = _stream()
        Lv = len(self.levels)
        lay = self.layers
        self._conv_back(lay["out"], b["dout"], b[f"y_dec0b"] if Lv > 1 else b["y_enc0b"], None, b["dx_out"], n, h, w)
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

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_count += 1
        _call("hpt_nn_adam", _stream(), _p(self.params), _p(self.grads), _p(self.adam_m), _p(self.adam_v),
              self.param_count, lr, beta1, beta2, eps, self.step_count)
        self.prepare_weights()

    def train_step(self, x: torch.Tensor, target_log: torch.Tensor, n: int, h: int, w: int, lr: float) -> torch.Tensor:
        """One step on a batch: forward, L1 loss in log space, backward, Adam.
        Returns the loss as a device tensor (read it when convenient)."""
        b = self._buffers(n, h, w, True)
        out = self.forward_input(x, n, h, w, train=True)
        if target_log.data_ptr() != b["target"].data_ptr():
            b["target"].copy_(target_log.reshape(-1))
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _call("hpt_nn_l1_loss", _stream(), _p(out), _p(b["target"]), _p(b["dout"]), _p(loss), n * h * w, OUT_PAD)
        self.backward(n, h, w)
        self.adam_step(lr)
        return loss

    # -------------------------------------------------------- inference
    def make_input(self, color, albedo, normal, x0: int, y0: int, h: int, w: int, out: torch.Tensor | None = None):
        """Network input (h*w*16 bf16) of the h x w window at (x0, y0) of
        float4 device images (H, W, 4)."""
        sh, sw = color.shape[:2]
        if out is None:
            out = torch.empty(h * w * IN_CHANNELS, dtype=torch.int16, device=self.device)
        _call("hpt_nn_make_input", _stream(), _p(color), _p(albedo), _p(normal), sw, sh, x0, y0, _p(out), h, w)
        return out

    def denoise(self, color: torch.Tensor, albedo: torch.Tensor, normal: torch.Tensor) -> torch.Tensor:
        """Denoised colour (H, W, 4) of float4 device images (H, W, 4); the
        network runs on the image padded up to a multiple of 16."""
        H, W = color.shape[:2]
        step = 1 << (len(self.levels) - 1)
        h, w = -(-H // step) * step, -(-W // step) * step
        b = self._buffers(1, h, w, False)
        self.make_input(color, albedo, normal, 0, 0, h, w, out=b["x"])
        out = self.forward_input(b["x"], 1, h, w)
        dst = torch.empty((H, W, 4), dtype=torch.float32, device=self.device)
        _call("hpt_nn_write_output", _stream(), _p(out), _p(dst), W, H, 0, 0, h, w)
        return dst

    def flops_forward(self, n: int, h: int, w: int) -> int:
        """Multiply-adds x 2 of the convolutions in one forward pass."""
        total = 0
        for name, L in self.layers.items():
            l = self._level(name)
            total += 2 * n * (h >> l) * (w >> l) * L.k * L.k * L.cin * L.cout
        return total


def read_renderer_images(r: "hpt.HalaRenderer", device: int = 0):
    """(colour, albedo, normal) float4 device tensors of a renderer's
    accumulated images."""
    dev = torch.device("cuda", device)
    return tuple(torch.from_numpy(np.ascontiguousarray(r.read_image(k))).to(dev)
                 for k in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL))


def weights_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_denoiser.npz")
