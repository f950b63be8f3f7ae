"""Denoiser kernels (include/hpt_nn.h) on the GPU against the numpy oracle
(oracle/unet_ref.py): each convolution pass within the float32 summation
bound, elementwise steps bit-exact, the whole network's output and gradients
within bf16 tolerance, determinism, and a training run that lowers the loss."""
import ctypes as C
import os

import numpy as np
import pytest

import unet_ref as U

pytestmark = pytest.mark.gpu

U24 = 2.0 ** -24


@pytest.fixture(scope="module")
def dn():
    import denoiser

    denoiser.nn_library()
    return denoiser


def _dn():
    import denoiser

    return denoiser


def _dev_bf16(a):
    """float32 numpy -> device buffer of the bf16 bit patterns."""
    bits = (U.bf16(a).view(np.uint32) >> 16).astype(np.uint16)
    return _dn().DeviceArray.from_numpy(bits.reshape(-1))


def _host_bf16(t, shape):
    bits = t.numpy().reshape(-1).view(np.uint16).astype(np.uint32) << 16
    return bits.view(np.float32).reshape(shape)


def _dev(a):
    return _dn().DeviceArray.from_numpy(np.ascontiguousarray(a, np.float32).reshape(-1))


def _zeros(count, dtype=np.float32):
    return _dn().DeviceArray(count, dtype)


def _p(t):
    return _dn()._p(t)


def _call(dn, name, *args):
    assert getattr(dn.nn_library(), name)(*args) == 0, name


def _bound(absx, absw, k, op):
    """|error| bound of a float32 sum of K products: (K - 1) u sum |terms|."""
    terms = op(absx, absw, k)
    K = absx.shape[-1] * k * k
    return 2 * (K + 1) * U24 * terms + 1e-30


CONV_SHAPES = [  # n, h, w, cin, cout, k
    (2, 8, 12, 16, 32, 3), (1, 16, 16, 224, 96, 3), (1, 8, 8, 80, 32, 3), (2, 16, 8, 32, 3, 1),
    (1, 6, 10, 48, 48, 3), (1, 5, 7, 112, 48, 3), (1, 4, 4, 128, 128, 3), (3, 7, 9, 160, 64, 3),
]


def _weights(dn, rng, cin, cout, k):
    cp = dn.pad16(cout)
    master = rng.normal(0, (2.0 / (k * k * cin)) ** 0.5, (cout, k * k, cin)).astype(np.float32)
    w = _zeros(cp * k * k * cin, np.uint16)
    wt = _zeros(cin * k * k * cp, np.uint16)
    _call(dn, "hpt_nn_prepare_weights", dn.stream(), _p(_dev(master)), _p(w), _p(wt), cin, cout, cp, k)
    dn.synchronize()
    return master, w, wt, cp


@pytest.mark.parametrize("shape", CONV_SHAPES)
def test_prepare_weights_layouts(dn, shape):
    _, _, _, cin, cout, k = shape
    rng = np.random.default_rng(1)
    master, w, wt, cp = _weights(dn, rng, cin, cout, k)
    wb = U.bf16(master)
    wpad = np.zeros((cp, k * k, cin), np.float32)
    wpad[:cout] = wb
    assert np.array_equal(_host_bf16(w, wpad.shape), wpad)
    # wt[ci][t'][co] = w[co][k*k - 1 - t'][ci]
    ref_t = wpad[:, ::-1, :].transpose(2, 1, 0)
    assert np.array_equal(_host_bf16(wt, (cin, k * k, cp)), ref_t)


@pytest.mark.parametrize("shape", CONV_SHAPES)
def test_conv_forward_matches_oracle(dn, shape):
    n, h, wd, cin, cout, k = shape
    print(f"conv forward {shape}", flush=True)
    rng = np.random.default_rng(2)
    x = U.bf16(rng.normal(0, 1, (n, h, wd, cin)).astype(np.float32))
    master, w, _, cp = _weights(dn, rng, cin, cout, k)
    b = rng.normal(0, 0.1, cout).astype(np.float32)
    y = _zeros(n * h * wd * cout, np.uint16)
    yf = _zeros(n * h * wd * cout)
    relu = int(cout > 3)
    _call(dn, "hpt_nn_conv_forward", dn.stream(), _p(_dev_bf16(x)), _p(w), _p(_dev(b)), _p(y), _p(yf), n, h, wd, cin, cout,
          cp, k, relu)
    dn.synchronize()
    wb = U.bf16(master).astype(np.float64)
    z = U.conv_forward(x.astype(np.float64), wb, b.astype(np.float64), k)
    ref = np.maximum(z, 0) if relu else z
    bound = _bound(np.abs(x).astype(np.float64), np.abs(wb), k,
                   lambda ax, aw, kk: U.conv_forward(ax, aw, np.abs(b).astype(np.float64), kk))
    got = yf.numpy().reshape(ref.shape)
    assert np.all(np.abs(got - ref) <= bound), float(np.max(np.abs(got - ref) - bound))
    # the bf16 output is the float32 value rounded to nearest even
    assert np.array_equal(_host_bf16(y, ref.shape), U.bf16(got))


# Images of at least 256 16x16 tiles take the LDS-tiled forward kernel (the
# 112-channel case needs 72 KB of LDS); the shapes above take the pixel-linear one.
TILED_SHAPES = [(4, 128, 128, 16, 32, 3), (4, 128, 128, 112, 48, 3), (1, 256, 272, 80, 32, 3)]


@pytest.mark.parametrize("shape", TILED_SHAPES)
def test_conv_forward_tiled_matches_oracle(dn, shape):
    test_conv_forward_matches_oracle(dn, shape)


UPCAT_SHAPES = [  # n, h, w, c_low, c_skip, cout: pixel-linear, tiled, tiled with 72 KB, linear (160 channels)
    (2, 16, 24, 32, 16, 32), (4, 128, 128, 48, 32, 32), (4, 128, 128, 64, 48, 48), (1, 64, 64, 96, 64, 64)]


@pytest.mark.parametrize("shape", UPCAT_SHAPES)
def test_conv_forward_upcat_equals_upsample_concat_conv(dn, shape):
    """The decoder's fused input (nearest 2x upsampling of `low` concatenated
    before `skip`) gives the bits of the three separate steps."""
    n, h, wd, cl, cs, cout = shape
    print(f"conv upcat {shape}", flush=True)
    rng = np.random.default_rng(7)
    low = _dev_bf16(rng.normal(0, 1, (n, h // 2, wd // 2, cl)).astype(np.float32))
    skip = _dev_bf16(rng.normal(0, 1, (n, h, wd, cs)).astype(np.float32))
    _, w, _, cp = _weights(dn, rng, cl + cs, cout, 3)
    b = _dev(rng.normal(0, 0.1, cout).astype(np.float32))
    up = _zeros(n * h * wd * cl, np.uint16)
    cat = _zeros(n * h * wd * (cl + cs), np.uint16)
    y_ref, y = _zeros(n * h * wd * cout, np.uint16), _zeros(n * h * wd * cout, np.uint16)
    _call(dn, "hpt_nn_upsample_forward", dn.stream(), _p(low), _p(up), n, h // 2, wd // 2, cl)
    _call(dn, "hpt_nn_concat_forward", dn.stream(), _p(up), _p(skip), _p(cat), n * h * wd, cl, cs)
    _call(dn, "hpt_nn_conv_forward", dn.stream(), _p(cat), _p(w), _p(b), _p(y_ref), None, n, h, wd, cl + cs, cout, cp, 3, 1)
    _call(dn, "hpt_nn_conv_forward_upcat", dn.stream(), _p(low), _p(skip), _p(w), _p(b), _p(y), None, n, h, wd, cl, cs, cout,
          cp, 1)
    dn.synchronize()
    assert np.array_equal(y.numpy(), y_ref.numpy())
    bad = dn.nn_library().hpt_nn_conv_forward_upcat(dn.stream(), _p(low), _p(skip), _p(w), _p(b), _p(y), None, n, h + 1,
                                                      wd, cl, cs, cout, cp, 1)
    assert bad != 0  # odd height: no half-resolution tensor to upsample


POOL_SHAPES = [  # n, h, w, cin, cout: tiled (cin 32, 48, 64, the encoders' second convolutions), pixel-linear
    (4, 128, 128, 32, 32), (4, 128, 128, 48, 48), (2, 256, 144, 64, 64), (1, 30, 40, 96, 96)]


@pytest.mark.parametrize("shape", POOL_SHAPES)
def test_conv_forward_pool_equals_conv_then_pool(dn, shape):
    """The encoder's second convolution with the next level's 2x2 average
    pooling written from the same pass gives the bits of the two calls."""
    n, h, wd, cin, cout = shape
    print(f"conv pool {shape}", flush=True)
    rng = np.random.default_rng(11)
    x = _dev_bf16(rng.normal(0, 1, (n, h, wd, cin)).astype(np.float32))
    _, w, _, cp = _weights(dn, rng, cin, cout, 3)
    b = _dev(rng.normal(0, 0.1, cout).astype(np.float32))
    y_ref, y = _zeros(n * h * wd * cout, np.uint16), _zeros(n * h * wd * cout, np.uint16)
    p_ref, pool = _zeros(n * h * wd * cout // 4, np.uint16), _zeros(n * h * wd * cout // 4, np.uint16)
    _call(dn, "hpt_nn_conv_forward", dn.stream(), _p(x), _p(w), _p(b), _p(y_ref), None, n, h, wd, cin, cout, cp, 3, 1)
    _call(dn, "hpt_nn_pool_forward", dn.stream(), _p(y_ref), _p(p_ref), n, h, wd, cout)
    _call(dn, "hpt_nn_conv_forward_pool", dn.stream(), _p(x), _p(w), _p(b), _p(y), _p(pool), n, h, wd, cin, cout, cp, 1)
    dn.synchronize()
    assert np.array_equal(y.numpy(), y_ref.numpy())
    assert np.array_equal(pool.numpy(), p_ref.numpy())
    bad = dn.nn_library().hpt_nn_conv_forward_pool(dn.stream(), _p(x), _p(w), _p(b), _p(y), _p(pool), n, h - 1, wd,
                                                     cin, cout, cp, 1)
    assert bad != 0  # odd height: no 2x2 pooling


@pytest.mark.parametrize("shape", CONV_SHAPES)
def test_conv_backward_data_matches_oracle(dn, shape):
    n, h, wd, cin, cout, k = shape
    print(f"conv backward data {shape}", flush=True)
    rng = np.random.default_rng(3)
    master, _, wt, cp = _weights(dn, rng, cin, cout, k)
    dy = np.zeros((n, h, wd, cp), np.float32)
    dy[..., :cout] = rng.normal(0, 1, (n, h, wd, cout))
    dx = _zeros(n * h * wd * cin)
    _call(dn, "hpt_nn_conv_backward_data", dn.stream(), _p(_dev(dy)), _p(wt), _p(dx), n, h, wd, cin, cout, cp, k)
    dn.synchronize()
    dyb = U.bf16(dy[..., :cout]).astype(np.float64)
    wb = U.bf16(master).astype(np.float64)
    ref = U.conv_backward_data(dyb, wb, k)
    bound = 2 * (cout * k * k + 1) * U24 * U.conv_backward_data(np.abs(dyb), np.abs(wb), k) + 1e-30
    got = dx.numpy().reshape(ref.shape)
    assert np.all(np.abs(got - ref) <= bound), float(np.max(np.abs(got - ref) - bound))


@pytest.mark.parametrize("shape", CONV_SHAPES)
def test_conv_backward_weights_matches_oracle(dn, shape):
    n, h, wd, cin, cout, k = shape
    print(f"conv backward weights {shape}", flush=True)
    rng = np.random.default_rng(4)
    cp = dn.pad16(cout)
    x = U.bf16(rng.normal(0, 1, (n, h, wd, cin)).astype(np.float32))
    dy = np.zeros((n, h, wd, cp), np.float32)
    dy[..., :cout] = rng.normal(0, 1, (n, h, wd, cout))
    lib = dn.nn_library()
    scratch = _zeros(lib.hpt_nn_wgrad_scratch_floats(n, h, wd, cin, cp, k))
    dw = _zeros(cout * k * k * cin)
    db = _zeros(cout)
    _call(dn, "hpt_nn_conv_backward_weights", dn.stream(), _p(_dev(dy)), _p(_dev_bf16(x)), _p(dw), _p(db), _p(scratch), n, h,
          wd, cin, cout, cp, k)
    dn.synchronize()
    dyb = U.bf16(dy[..., :cout]).astype(np.float64)
    ref_w, _ = U.conv_backward_weights(dyb, x.astype(np.float64), k)
    P = n * h * wd
    bound_w = 2 * (P + 1) * U24 * U.conv_backward_weights(np.abs(dyb), np.abs(x).astype(np.float64), k)[0] + 1e-30
    got_w = dw.numpy().reshape(ref_w.shape)
    assert np.all(np.abs(got_w - ref_w) <= bound_w), float(np.max(np.abs(got_w - ref_w) - bound_w))
    dy64 = dy[..., :cout].astype(np.float64).reshape(-1, cout)
    ref_b = dy64.sum(0)
    bound_b = 2 * (P + 1) * U24 * np.abs(dy64).sum(0)
    assert np.all(np.abs(db.numpy() - ref_b) <= bound_b)


def test_elementwise_steps_bit_exact(dn):
    rng = np.random.default_rng(5)
    n, h, w, c = 2, 6, 10, 48
    x = U.bf16(rng.normal(0, 1, (n, h, w, c)).astype(np.float32))
    xd = _dev_bf16(x)
    st = None
    y = _zeros(n * h * w * c // 4, np.uint16)
    _call(dn, "hpt_nn_pool_forward", st, _p(xd), _p(y), n, h, w, c)
    assert np.array_equal(_host_bf16(y, (n, h // 2, w // 2, c)), U.bf16(U.pool_forward(x)))
    g = rng.normal(0, 1, (n, h // 2, w // 2, c)).astype(np.float32)
    dx = _zeros(n * h * w * c)
    _call(dn, "hpt_nn_pool_backward", st, _p(_dev(g)), _p(dx), n, h, w, c)
    assert np.array_equal(dx.numpy().reshape(n, h, w, c), U.pool_backward(g))
    up = _zeros(n * h * w * c * 4, np.uint16)
    _call(dn, "hpt_nn_upsample_forward", st, _p(xd), _p(up), n, h, w, c)
    assert np.array_equal(_host_bf16(up, (n, 2 * h, 2 * w, c)), U.upsample_forward(x))
    g2 = rng.normal(0, 1, (n, 2 * h, 2 * w, c)).astype(np.float32)
    _call(dn, "hpt_nn_upsample_backward", st, _p(_dev(g2)), _p(dx), n, h, w, c)
    assert np.array_equal(dx.numpy().reshape(n, h, w, c), U.upsample_backward(g2))
    # concat and its split
    b2 = U.bf16(rng.normal(0, 1, (n, h, w, 32)).astype(np.float32))
    cat = _zeros(n * h * w * 80, np.uint16)
    _call(dn, "hpt_nn_concat_forward", st, _p(xd), _p(_dev_bf16(b2)), _p(cat), n * h * w, c, 32)
    ref_cat = np.concatenate([x, b2], axis=-1)
    assert np.array_equal(_host_bf16(cat, ref_cat.shape), ref_cat)
    gc = rng.normal(0, 1, ref_cat.shape).astype(np.float32)
    da = _zeros(n * h * w * c)
    dbb = _zeros(n * h * w * 32)
    _call(dn, "hpt_nn_concat_backward", st, _p(_dev(gc)), _p(da), _p(dbb), n * h * w, c, 32)
    assert np.array_equal(da.numpy().reshape(x.shape), gc[..., :c])
    assert np.array_equal(dbb.numpy().reshape(b2.shape), gc[..., c:])
    # relu derivative from the stored bf16 output, accumulate
    yr = U.bf16(np.maximum(rng.normal(0, 1, 1000), 0).astype(np.float32))
    yr[::7] = -0.0
    gr = rng.normal(0, 1, 1000).astype(np.float32)
    gd = _dev(gr)
    _call(dn, "hpt_nn_relu_backward", st, _p(gd), _p(_dev_bf16(yr)), 1000)
    assert np.array_equal(gd.numpy(), gr * (yr > 0))
    acc = _dev(gr)
    _call(dn, "hpt_nn_accumulate", st, _p(acc), _p(_dev(yr)), 1000)
    assert np.array_equal(acc.numpy(), gr + yr)
    # the residual path adds the input's bf16 log colour to the output
    P = 77
    xin = U.bf16(rng.normal(0, 1, (P, 16)).astype(np.float32))
    o = rng.normal(0, 1, (P, 3)).astype(np.float32)
    od = _dev(o)
    _call(dn, "hpt_nn_add_residual", st, _p(od), _p(_dev_bf16(xin)), P, 16)
    assert np.array_equal(od.numpy().reshape(P, 3), o + xin[:, :3])


def test_make_input_loss_adam_and_output(dn):
    rng = np.random.default_rng(6)
    H, W = 9, 13
    color = np.abs(rng.normal(0, 2, (H, W, 4))).astype(np.float32)
    color[0, 0, 0] = -0.5  # negative radiance is clamped to 0
    albedo = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    normal = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    h, w, x0, y0 = 8, 16, 2, 3  # the window sticks out of the image on the right and bottom
    x = _zeros(h * w * 16, np.uint16)
    _call(dn, "hpt_nn_make_input", dn.stream(), _p(_dev(color)), _p(_dev(albedo)), _p(_dev(normal)), W, H, x0, y0, _p(x), h,
          w, 0, 0.625)
    got = _host_bf16(x, (h, w, 16))
    full = U.make_input(color, albedo, normal, 0.625)
    ref = np.zeros((h, w, 16), np.float32)
    ys, xs = slice(y0, min(H, y0 + h)), slice(x0, min(W, x0 + w))
    ref[:ys.stop - y0, :xs.stop - x0] = full[ys, xs]
    ulp = np.abs(got.view(np.int32) - U.bf16(ref).view(np.int32))
    assert ulp.max() <= 1 << 16  # log1p may differ by an ulp before the bf16 rounding
    # mirrored windows read the same pixels in reverse order; the target is log(1 + colour)
    for m in (1, 2, 3):
        xm = _zeros(h * w * 16, np.uint16)
        _call(dn, "hpt_nn_make_input", dn.stream(), _p(_dev(color)), _p(_dev(albedo)), _p(_dev(normal)), W, H, x0, y0,
              _p(xm), h, w, m, 0.625)
        gm = _host_bf16(xm, (h, w, 16))
        exp = got[:, ::-1] if m == 1 else (got[::-1] if m == 2 else got[::-1, ::-1])
        assert np.array_equal(gm, exp)
    tg = _zeros(h * w * 3)
    _call(dn, "hpt_nn_make_target", dn.stream(), _p(_dev(color)), W, H, x0, y0, _p(tg), h, w, 0)
    ref_t = np.zeros((h, w, 3), np.float32)
    ref_t[:ys.stop - y0, :xs.stop - x0] = np.log1p(np.maximum(color[ys, xs, :3], 0))
    assert np.allclose(tg.numpy().reshape(h, w, 3), ref_t, rtol=5e-7, atol=1e-7)

    P = 300
    out = rng.normal(0, 1, (P, 3)).astype(np.float32)
    tgt = rng.normal(0, 1, (P, 3)).astype(np.float32)
    dy = _zeros(P * 16)
    dy.fill(7.0)
    loss = _zeros(1)
    _call(dn, "hpt_nn_l1_loss", dn.stream(), _p(_dev(out)), _p(_dev(tgt)), _p(dy), _p(loss), P, 16)
    ref_loss, ref_g = U.l1_loss(out.astype(np.float64), tgt.astype(np.float64))
    gd = dy.numpy().reshape(P, 16)
    assert np.array_equal(gd[:, 3:], np.zeros((P, 13), np.float32))
    assert np.allclose(gd[:, :3], ref_g, rtol=1e-6, atol=0)
    assert float(loss.numpy()[0]) == pytest.approx(ref_loss, rel=1e-5)

    p = rng.normal(0, 1, 500).astype(np.float32)
    g = rng.normal(0, 1, 500).astype(np.float32)
    pd, md, vd = _dev(p), _zeros(500), _zeros(500)
    rp, rm, rv = p.astype(np.float64), np.zeros(500), np.zeros(500)
    for t in (1, 2, 3):
        gt = g * t
        _call(dn, "hpt_nn_adam", dn.stream(), _p(pd), _p(_dev(gt)), _p(md), _p(vd), 500, 1e-2, 0.9, 0.999, 1e-8, t)
        rp, rm, rv = U.adam(rp, gt.astype(np.float64), rm, rv, 1e-2, 0.9, 0.999, 1e-8, t)
    assert np.allclose(pd.numpy(), rp, rtol=1e-5, atol=1e-6)

    o = rng.normal(0, 1, (h * w, 3)).astype(np.float32)
    dst = _zeros(H * W * 4)
    dst.fill(-3.0)
    _call(dn, "hpt_nn_write_output", dn.stream(), _p(_dev(o)), _p(dst), W, H, x0, y0, h, w)
    img = dst.numpy().reshape(H, W, 4)
    o3 = o.reshape(h, w, 3)
    win = img[ys, xs]
    assert np.allclose(win[..., :3], np.maximum(np.expm1(o3[:ys.stop - y0, :xs.stop - x0]), 0), rtol=1e-5, atol=1e-6)
    assert np.all(win[..., 3] == 1.0) and np.all(img[:y0] == -3.0)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_unet_forward_backward_match_oracle(dn):
    n, h, w = 2, 32, 48
    print("unet full network vs oracle", flush=True)
    net = dn.UNet(seed=11)
    rng = np.random.default_rng(12)
    params = {k: (wt, b + rng.normal(0, 0.02, b.shape).astype(np.float32)) for k, (wt, b) in net.get_params().items()}
    net.set_params(params)
    x = U.bf16(np.concatenate([rng.uniform(0, 1, (n, h, w, 9)), np.zeros((n, h, w, 7))], -1).astype(np.float32))
    tgt = rng.normal(0, 0.5, (n, h, w, 3)).astype(np.float32)
    out = net.forward_input(_dev_bf16(x), n, h, w, train=True)
    b = net._buffers(n, h, w, True)
    b["target"].upload(tgt.reshape(-1))
    loss = _zeros(1)
    _call(dn, "hpt_nn_l1_loss", dn.stream(), _p(out), _p(b["target"]), _p(b["dout"]), _p(loss), n * h * w, 16)
    net.backward(n, h, w)
    got_out = out.numpy().reshape(n, h, w, 3)
    grads = net.get_grads()

    ref = U.UNet(params, bf16_rounding=True)
    r_out, cache = ref.forward(x)
    assert _rel(got_out, r_out) < 1e-2
    dout = b["dout"].numpy().reshape(n, h, w, 16)[..., :3]
    r_grads = ref.backward(dout, cache)
    for name, (dw, db) in r_grads.items():
        assert _rel(grads[name][0], dw) < 3e-2, (name, _rel(grads[name][0], dw))
        assert _rel(grads[name][1], db) < 3e-2, (name, _rel(grads[name][1], db))

    # the backward pass is deterministic: a second run gives the same bits
    first = net.grads.numpy()
    net.forward_input(_dev_bf16(x), n, h, w, train=True)
    net.backward(n, h, w)
    assert np.array_equal(first, net.grads.numpy())


def test_training_lowers_the_loss(dn):
    """A few hundred Adam steps on a synthetic denoising task: smooth colour
    fields plus noise whose variance the albedo channel does not explain."""
    print("unet training on a synthetic task", flush=True)
    n, h, w = 4, 64, 64
    rng = np.random.default_rng(13)
    net = dn.UNet(seed=14)
    yy, xx = np.mgrid[0:h, 0:w] / 64.0

    def batch():
        clean = np.stack([np.stack([0.5 + 0.4 * np.sin(6 * xx * rng.uniform(0.5, 1.5) + rng.uniform(0, 6)),
                                    0.5 + 0.4 * np.cos(5 * yy * rng.uniform(0.5, 1.5) + rng.uniform(0, 6)),
                                    0.5 + 0.3 * np.sin(4 * (xx + yy) + rng.uniform(0, 6))], -1) for _ in range(n)])
        noisy = np.maximum(clean + rng.normal(0, 0.3, clean.shape), 0)
        x = np.zeros((n, h, w, 16), np.float32)
        x[..., 0:3] = np.log1p(noisy)
        x[..., 3:6] = 0.5
        x[..., 6:9] = 0.5
        return _dev_bf16(x), _dev(np.log1p(clean))

    losses = []
    for it in range(200):
        x, t = batch()
        losses.append(net.train_step(x, t, n, h, w, lr=2e-3))
    vals = [float(v) for v in losses]
    first, last = np.mean(vals[:10]), np.mean(vals[-10:])
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert np.all(np.isfinite(vals))
    assert last < 0.5 * first


def test_trained_network_beats_the_plain_estimate(dn):
    """The committed network (trained on the GPU by the kernels above, on
    4200 random room and outdoor scenes: no displaced icosphere, the layout of
    the large-mesh benchmark scene) against the plain 16-spp estimate on
    held-out scenes of all three layouts: relMSE to a 2048-spp reference from
    disjoint frames."""
    import denoise_train

    path = dn.weights_path()
    assert os.path.exists(path), f"{path} is missing (hala-pathtracer_amd/denoise_train.py writes it)"
    net = dn.UNet.load(path)
    gains = []
    for seed in (900, 901, 902):
        print(f"held-out scene {seed}", flush=True)
        s = denoise_train.render_sample(seed, 256, 256, spp_levels=(16,), ref_spp=2048)
        ref = s.reference.numpy()
        plain = denoise_train.relmse(s.noisy[16][0].numpy(), ref)
        den = denoise_train.relmse(net.denoise(s.noisy[16], 256, 256, 16).numpy(), ref)
        print(f"  relMSE plain {plain:.5f} denoised {den:.5f}", flush=True)
        gains.append(plain / den)
    assert min(gains) > 2.0 and float(np.exp(np.mean(np.log(gains)))) > 4.0, gains


def test_denoise_cli_replaces_oidn_on_saved_images(dn, tmp_path):
    """denoise.py's flow (find *_color.pfm with _albedo/_normal, write
    _final.pfm) through the learned denoiser: the written image is the
    network's output for those images, and it is closer to a reference than
    the noisy colour."""
    import denoise_cli
    import denoise_train
    import hpt

    out = tmp_path / "out" / "scene"
    out.mkdir(parents=True)
    s = denoise_train.render_sample(901, 128, 128, spp_levels=(16,), ref_spp=1024)
    prefix = str(out / "cornellish")
    for img, suffix in zip(s.noisy[16], ("color", "albedo", "normal")):
        hpt.write_pfm(f"{prefix}_{suffix}.pfm", img.numpy()[..., :3])
    with open(prefix + "_spp.txt", "w") as f:
        f.write("16\n")
    (tmp_path / "out" / "lonely_color.pfm").write_bytes(b"")  # no albedo/normal: skipped
    assert denoise_cli.main(["--dir", str(tmp_path / "out")]) == 0
    final = hpt.read_pfm(prefix + "_final.pfm")
    net = dn.UNet.load(dn.weights_path())
    direct = net.denoise(s.noisy[16], 128, 128, 16).numpy()[..., :3]
    assert np.allclose(final, direct, rtol=1e-6, atol=1e-6)
    ref = s.reference.numpy()
    assert denoise_train.relmse(final, ref) < denoise_train.relmse(s.noisy[16][0].numpy(), ref)
    assert not (tmp_path / "out" / "lonely_final.pfm").exists()
