"""Pins of the denoiser oracle (oracle/unet_ref.py) by mathematics and an
independent library: bf16 rounding against torch's bfloat16 cast, the
convolution against a brute-force loop, every backward against central
finite differences (float64) and its adjoint identity, the whole network and
its gradients against torch autograd (float64), Adam's first step against
its closed form."""
import numpy as np
import pytest
import torch

import unet_ref as U


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_bf16_matches_torch_round_to_nearest_even():
    rng = _rng(1)
    a = np.concatenate([rng.normal(0, 10, 4096).astype(np.float32),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0, 3.0e38, 1e-39])])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16(a).view(np.uint32), ref.view(np.uint32))
    # ties go to the even mantissa
    assert U.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert U.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == np.float32(1 + 2 ** -6)


def _brute_conv(x, w, b, k):
    n, h, wd, cin = x.shape
    cout = w.shape[0]
    r = k // 2
    z = np.zeros((n, h, wd, cout))
    for i in range(n):
        for y in range(h):
            for xx in range(wd):
                for co in range(cout):
                    s = b[co]
                    for dy in range(k):
                        for dx in range(k):
                            yy, xs = y + dy - r, xx + dx - r
                            if 0 <= yy < h and 0 <= xs < wd:
                                s += np.dot(x[i, yy, xs], w[co, dy * k + dx])
                    z[i, y, xx, co] = s
    return z


@pytest.mark.parametrize("k", [1, 3])
def test_conv_forward_matches_brute_force(k):
    rng = _rng(2)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(4, k * k, 3))
    b = rng.normal(size=4)
    assert np.allclose(U.conv_forward(x, w, b, k), _brute_conv(x, w, b, k), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 3])
def test_conv_backward_data_is_the_adjoint(k):
    rng = _rng(3)
    x = rng.normal(size=(2, 4, 5, 3))
    w = rng.normal(size=(6, k * k, 3))
    y = rng.normal(size=(2, 4, 5, 6))
    lhs = np.sum(U.conv_forward(x, w, np.zeros(6), k) * y)
    rhs = np.sum(x * U.conv_backward_data(y, w, k))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_conv_backward_weights_by_finite_differences():
    rng = _rng(4)
    k = 3
    x = rng.normal(size=(2, 4, 4, 3))
    w = rng.normal(size=(2, 9, 3))
    b = rng.normal(size=2)
    g = rng.normal(size=(2, 4, 4, 2))  # loss = sum(g * conv(x))
    dw, db = U.conv_backward_weights(g, x, k)
    eps = 1e-6
    for idx in [(0, 0, 0), (1, 4, 2), (0, 8, 1), (1, 2, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = (np.sum(g * U.conv_forward(x, wp, b, k)) - np.sum(g * U.conv_forward(x, wm, b, k))) / (2 * eps)
        assert abs(fd - dw[idx]) < 1e-6 * max(1.0, abs(fd))
    assert np.allclose(db, g.reshape(-1, 2).sum(0))


def test_pool_and_upsample_adjoints_and_values():
    rng = _rng(5)
    x = rng.normal(size=(2, 4, 6, 3))
    y = rng.normal(size=(2, 2, 3, 3))
    p = U.pool_forward(x)
    assert np.allclose(p[1, 1, 2], x[1, 2:4, 4:6].mean(axis=(0, 1)))
    assert abs(np.sum(p * y) - np.sum(x * U.pool_backward(y))) < 1e-12
    u = U.upsample_forward(y)
    assert np.array_equal(u[0, 3, 5], y[0, 1, 2])
    assert abs(np.sum(u * x) - np.sum(y * U.upsample_backward(x))) < 1e-12


def test_l1_loss_gradient_by_finite_differences():
    rng = _rng(6)
    out = rng.normal(size=(3, 4, 3))
    t = out + rng.choice([-1.0, 1.0], size=out.shape) * rng.uniform(0.1, 1.0, size=out.shape)
    loss, g = U.l1_loss(out, t)
    eps = 1e-7
    for idx in [(0, 0, 0), (2, 3, 1), (1, 2, 2)]:
        op, om = out.copy(), out.copy()
        op[idx] += eps
        om[idx] -= eps
        fd = (U.l1_loss(op, t)[0] - U.l1_loss(om, t)[0]) / (2 * eps)
        assert abs(fd - g[idx]) < 1e-7
    assert loss == pytest.approx(np.abs(out - t).mean())


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 2.0, 3.0, 4.0])
    p1, m, v = U.adam(p, g, np.zeros(4), np.zeros(4), lr=0.1, b1=0.9, b2=0.999, eps=1e-8, t=1)
    # bias-corrected m = g, v = g^2 after one step: the update is lr * g / (|g| + eps)
    assert np.allclose(p1, p - 0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def _torch_unet(params, levels, x):
    """The same network written with torch.nn.functional (float64)."""
    import torch.nn.functional as F

    tp = {}
    for name, (w, b) in params.items():
        cout, kk, cin = w.shape
        k = int(round(kk ** 0.5))
        wt = torch.tensor(w.reshape(cout, k, k, cin).transpose(0, 3, 1, 2).copy(), requires_grad=True)
        tp[name] = (wt, torch.tensor(b.copy(), requires_grad=True), k)

    def conv(name, t, relu=True):
        w, b, k = tp[name]
        z = F.conv2d(t, w, b, padding=k // 2)
        return F.relu(z) if relu else z

    t = torch.tensor(x.transpose(0, 3, 1, 2).copy())
    skips = []
    cur = t
    for l in range(len(levels)):
        if l > 0:
            cur = F.avg_pool2d(skips[-1], 2)
        cur = conv(f"enc{l}b", conv(f"enc{l}a", cur))
        skips.append(cur)
    for l in range(len(levels) - 2, -1, -1):
        u = F.interpolate(cur, scale_factor=2, mode="nearest")
        cur = conv(f"dec{l}b", conv(f"dec{l}a", torch.cat([u, skips[l]], dim=1)))
    out = conv("out", cur, relu=False) + t[:, :3]
    return out, tp


def test_unet_forward_and_gradients_match_torch_autograd():
    levels = (4, 6, 8)
    rng = _rng(7)
    params = {k: (w.astype(np.float64), b.astype(np.float64) + rng.normal(0, 0.05, b.shape))
              for k, (w, b) in U.init_params(levels, seed=3).items()}
    x = rng.normal(size=(2, 8, 12, U.IN_CHANNELS))
    net = U.UNet(params, levels, bf16_rounding=False)
    out, cache = net.forward(x)
    g = rng.normal(size=out.shape)  # loss = sum(g * out)
    grads = net.backward(g, cache)

    tout, tp = _torch_unet(params, levels, x)
    assert np.allclose(out, tout.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-10, atol=1e-12)
    (tout * torch.tensor(g.transpose(0, 3, 1, 2).copy())).sum().backward()
    for name, (dw, db) in grads.items():
        wt, bt, k = tp[name]
        tw = wt.grad.numpy().transpose(0, 2, 3, 1).reshape(dw.shape)
        assert np.allclose(dw, tw, rtol=1e-9, atol=1e-11), name
        assert np.allclose(db, bt.grad.numpy(), rtol=1e-9, atol=1e-11), name


def test_unet_gradient_by_finite_differences():
    levels = (4, 6)
    rng = _rng(8)
    params = {k: (w.astype(np.float64), b.astype(np.float64)) for k, (w, b) in U.init_params(levels, seed=5).items()}
    x = rng.normal(size=(1, 4, 6, U.IN_CHANNELS))
    t = rng.normal(size=(1, 4, 6, 3))

    def loss(p):
        out, _ = U.UNet(p, levels, bf16_rounding=False).forward(x)
        return 0.5 * np.sum((out - t) ** 2)

    net = U.UNet(params, levels, bf16_rounding=False)
    out, cache = net.forward(x)
    grads = net.backward(out - t, cache)
    eps = 1e-6
    for name, idx in [("enc0a", (1, 4, 3)), ("enc1b", (2, 0, 5)), ("dec0a", (0, 7, 9)), ("out", (2, 0, 1))]:
        w, b = params[name]
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = (loss({**params, name: (wp, b)}) - loss({**params, name: (wm, b)})) / (2 * eps)
        assert abs(fd - grads[name][0][idx]) < 1e-6 * max(1.0, abs(fd)), name


def test_layer_specs_channel_plan():
    specs = {s[0]: s for s in U.layer_specs()}
    assert specs["enc0a"][1:3] == (16, 32) and specs["enc4b"][1:3] == (128, 128)
    assert specs["dec3a"][1:3] == (224, 96) and specs["dec0a"][1:3] == (80, 32)
    assert specs["out"][1:4] == (32, 3, 1)
    n = sum(co * k * k * ci + co for _, ci, co, k, _ in U.layer_specs())
    assert 900_000 < n < 1_300_000
