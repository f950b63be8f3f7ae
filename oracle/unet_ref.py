"""CPU restatement of the learned denoiser (numpy): the U-Net, its backward
pass and its training step, for checking the HIP kernels of include/hpt_nn.h.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.

The reference has no counterpart (its denoise.py:7-30 shells out to Intel
OIDN), so nothing here follows a reference file; every function is pinned by
mathematics in tests/test_oracle_unet.py instead: convolutions against a
brute-force loop, every backward against central finite differences in
float64 and against its adjoint identity <A x, y> = <x, A^T y>, the loss
gradient by finite differences, Adam's first step by its closed form.

Two precisions.  With `bf16=True` the functions round where the GPU rounds:
weights and stored activations to bfloat16 (round to nearest even), and the
float32 gradient that enters a convolution's data- or weight-gradient
product to bfloat16, with float32 sums.  With `bf16=False` everything runs in
the input dtype (float64 for finite-difference checks).

Tensors are NHWC; a convolution weight is [cout][k*k][cin] with tap
t = 3 * dy + dx over the offsets (dy - k//2, dx - k//2), as on the GPU.
"""
from __future__ import annotations

import numpy as np

LEVELS = (32, 48, 64, 96, 128)
IN_CHANNELS = 16     # log(1 + colour), albedo, normal * 0.5 + 0.5, sample-count feature, zeros
OUT_CHANNELS = 3


def bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def _r(a, on):
    return bf16(a) if on else a


# ------------------------------------------------------------------ layers
def _taps(k):
    r = k // 2
    return [(dy - r, dx - r) for dy in range(k) for dx in range(k)]


def _shift(x, oy, ox):
    """s[n, y, x] = x[n, y + oy, x + ox] inside the image, zero outside."""
    n, h, w, c = x.shape
    s = np.zeros_like(x)
    ys, ye = max(0, -oy), min(h, h - oy)
    xs, xe = max(0, -ox), min(w, w - ox)
    if ys < ye and xs < xe:
        s[:, ys:ye, xs:xe] = x[:, ys + oy:ye + oy, xs + ox:xe + ox]
    return s


def conv_forward(x, w, b, k):
    """z[p][co] = sum_t sum_ci x[p + off(t)][ci] * w[co][t][ci] + b[co]."""
    n, h, wd, cin = x.shape
    z = np.zeros((n, h, wd, w.shape[0]), np.result_type(x, w))
    for t, (oy, ox) in enumerate(_taps(k)):
        z += _shift(x, oy, ox) @ w[:, t, :].T
    return z + b


def conv_backward_data(dz, w, k):
    """dx = the adjoint of conv_forward's linear part, applied to dz."""
    n, h, wd, _ = dz.shape
    dx = np.zeros((n, h, wd, w.shape[2]), np.result_type(dz, w))
    for t, (oy, ox) in enumerate(_taps(k)):
        dx += _shift(dz @ w[:, t, :], -oy, -ox)
    return dx


def conv_backward_weights(dz, x, k):
    """dw[co][t][ci] = sum_p dz[p][co] x[p + off(t)][ci]; db[co] = sum_p dz[p][co]."""
    cout, cin = dz.shape[-1], x.shape[-1]
    dz2 = dz.reshape(-1, cout)
    dw = np.zeros((cout, k * k, cin), np.result_type(dz, x))
    for t, (oy, ox) in enumerate(_taps(k)):
        dw[:, t, :] = dz2.T @ _shift(x, oy, ox).reshape(-1, cin)
    return dw, dz2.sum(axis=0)


def pool_forward(x):
    n, h, w, c = x.shape
    v = x.reshape(n, h // 2, 2, w // 2, 2, c)
    return 0.25 * ((v[:, :, 0, :, 0] + v[:, :, 0, :, 1]) + (v[:, :, 1, :, 0] + v[:, :, 1, :, 1]))


def pool_backward(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def upsample_forward(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample_backward(dy):
    n, h, w, c = dy.shape
    v = dy.reshape(n, h // 2, 2, w // 2, 2, c)
    return (v[:, :, 0, :, 0] + v[:, :, 0, :, 1]) + (v[:, :, 1, :, 0] + v[:, :, 1, :, 1])


def make_input(color, albedo, normal, aux=0.0):
    """(H, W, 4) float images -> (H, W, 16) network input."""
    h, w = color.shape[:2]
    x = np.zeros((h, w, IN_CHANNELS), np.float32)
    x[..., 0:3] = np.log1p(np.maximum(color[..., :3], 0).astype(np.float32))
    x[..., 3:6] = albedo[..., :3]
    x[..., 6:9] = normal[..., :3] * np.float32(0.5) + np.float32(0.5)
    x[..., 9] = aux
    return x


def l1_loss(out, target_log):
    """mean |out - target| and its gradient."""
    d = out - target_log
    return float(np.abs(d).mean()), np.sign(d) / d.size


def adam(p, g, m, v, lr, b1, b2, eps, t):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------------ network
def layer_specs(levels=LEVELS):
    """[(name, cin, cout, k, relu)] in forward order: two 3x3 convs per
    encoder level (2x2 average pooling between levels), two per decoder level
    after nearest 2x upsampling and concatenation [upsampled, skip], a final
    1x1 conv to RGB without activation."""
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


def init_params(levels=LEVELS, seed=0, dtype=np.float32):
    """He-normal weights [cout][k*k][cin], biases 0.01."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, k, _ in layer_specs(levels):
        std = np.sqrt(2.0 / (k * k * cin))
        params[name] = (rng.normal(0.0, std, (cout, k * k, cin)).astype(dtype), np.full(cout, 0.01, dtype))
    return params


class UNet:
    def __init__(self, params, levels=LEVELS, bf16_rounding=True, residual=True):
        self.params = params
        self.levels = tuple(levels)
        self.rb = bf16_rounding
        self.residual = residual
        self.specs = {s[0]: s for s in layer_specs(levels)}

    def _conv(self, name, x, cache):
        _, _, cout, k, relu = self.specs[name]
        w, b = self.params[name]
        z = conv_forward(x, _r(w, self.rb), b, k)
        if relu:
            y = _r(np.maximum(z, 0), self.rb)
        else:
            y = z  # the network output stays float32
        cache[name] = (x, y)
        return y

    def forward(self, x):
        """x (N, H, W, 16) with H, W multiples of 2^(levels-1); returns the
        output (N, H, W, 3) and the cache for backward."""
        cache = {}
        L = len(self.levels)
        skips = []
        cur = _r(x, self.rb)
        for l in range(L):
            if l > 0:
                cur = _r(pool_forward(skips[-1]), self.rb)
            cur = self._conv(f"enc{l}a", cur, cache)
            cur = self._conv(f"enc{l}b", cur, cache)
            skips.append(cur)
        for l in range(L - 2, -1, -1):
            u = upsample_forward(cur)
            cur = self._conv(f"dec{l}a", np.concatenate([u, skips[l]], axis=-1), cache)
            cur = self._conv(f"dec{l}b", cur, cache)
        out = self._conv("out", cur, cache)
        if self.residual:  # the output corrects the input's log colour
            out = out + _r(x, self.rb)[..., :OUT_CHANNELS]
        return out, cache

    def _conv_back(self, name, dy, cache, grads, need_dx=True):
        _, _, cout, k, relu = self.specs[name]
        w, _ = self.params[name]
        x, y = cache[name]
        dz = dy * (y > 0) if relu else dy
        dzr = _r(dz.astype(np.float32) if self.rb else dz, self.rb)
        grads[name] = conv_backward_weights(dzr, x, k)
        return conv_backward_data(dzr, _r(w, self.rb), k) if need_dx else None

    def backward(self, dout, cache):
        """Gradients {name: (dw, db)} of the loss whose gradient w.r.t. the
        output is dout."""
        grads = {}
        L = len(self.levels)
        d = self._conv_back("out", dout, cache, grads)
        dskip = [None] * L
        for l in range(L - 1):
            d = self._conv_back(f"dec{l}b", d, cache, grads)
            d = self._conv_back(f"dec{l}a", d, cache, grads)
            cu = self.levels[l + 1]
            du, dskip[l] = d[..., :cu], d[..., cu:]
            d = upsample_backward(du)  # gradient w.r.t. the level l+1 output
        for l in range(L - 1, -1, -1):
            if l < L - 1:
                d = dskip[l] + pool_backward(dpool)
            d = self._conv_back(f"enc{l}b", d, cache, grads)
            dpool = self._conv_back(f"enc{l}a", d, cache, grads, need_dx=l > 0)
        return grads
