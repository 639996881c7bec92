"""Pins of the FP64 U-Net oracle (oracle/unet_oracle.py).

Each layer is checked against torch's CPU float64 functional ops (an
independent implementation of convolution, transposed convolution and
average pooling), the whole network's loss gradient against torch autograd
on a network composed from those ops, the analytic gradient against central
finite differences, and the bfloat16 rounding against torch's float32 ->
bfloat16 conversion.  A transposed weight, a flipped tap, a wrong pixel
phase of the transposed convolution or a dropped skip gradient fails one of
them.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_oracle as U


def _nchw(x):
    return torch.from_numpy(np.ascontiguousarray(np.transpose(x, (0, 3, 1, 2))))


def _nhwc(t):
    return np.transpose(t.detach().numpy(), (0, 2, 3, 1))


def _torch_conv3_weight(W):
    co, _, ci = W.shape
    return torch.from_numpy(np.ascontiguousarray(W.reshape(co, 3, 3, ci).transpose(0, 3, 1, 2)))


def _torch_up2_weight(W):
    _, co, ci = W.shape
    return torch.from_numpy(np.ascontiguousarray(W.reshape(2, 2, co, ci).transpose(3, 2, 0, 1)))


def test_layers_match_torch():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 5, 3))
    W = rng.standard_normal((4, 9, 3))
    b = rng.standard_normal(4)
    want = F.conv2d(_nchw(x), _torch_conv3_weight(W), torch.from_numpy(b), padding=1)
    np.testing.assert_allclose(U.conv3(x, W, b), _nhwc(want), rtol=1e-12, atol=1e-12)
    Wu = rng.standard_normal((4, 5, 3))
    bu = rng.standard_normal(5)
    want = F.conv_transpose2d(_nchw(x), _torch_up2_weight(Wu), torch.from_numpy(bu), stride=2)
    np.testing.assert_allclose(U.up2(x, Wu, bu), _nhwc(want), rtol=1e-12, atol=1e-12)
    x4 = rng.standard_normal((2, 6, 4, 3))
    np.testing.assert_allclose(U.pool2(x4), _nhwc(F.avg_pool2d(_nchw(x4), 2)), rtol=1e-14)


def test_layer_backward_matches_torch():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 4, 3))
    W = rng.standard_normal((5, 9, 3))
    dy = rng.standard_normal((2, 6, 4, 5))
    xt = _nchw(x).requires_grad_()
    wt = _torch_conv3_weight(W).requires_grad_()
    bt = torch.zeros(5, dtype=torch.float64, requires_grad=True)
    F.conv2d(xt, wt, bt, padding=1).backward(_nchw(dy))
    dx, dW, db = U.conv3_backward(x, W, dy)
    np.testing.assert_allclose(dx, _nhwc(xt.grad), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dW, wt.grad.numpy().transpose(0, 2, 3, 1).reshape(5, 9, 3), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12)
    Wu = rng.standard_normal((4, 5, 3))
    dyu = rng.standard_normal((2, 12, 8, 5))
    xt = _nchw(x).requires_grad_()
    wt = _torch_up2_weight(Wu).requires_grad_()
    F.conv_transpose2d(xt, wt, None, stride=2).backward(_nchw(dyu))
    dx, dWu, _ = U.up2_backward(x, Wu, dyu)
    np.testing.assert_allclose(dx, _nhwc(xt.grad), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dWu, wt.grad.numpy().transpose(2, 3, 1, 0).reshape(4, 5, 3), rtol=1e-12, atol=1e-12)


def _torch_net(flat, x, kw):
    """The surrogate composed from torch ops, independently of the oracle."""
    P = U.unpack(flat, **kw)
    L = kw["levels"]
    a = x
    skips = []
    for l in range(L):
        for s in "ab":
            W, b = P[f"enc{l}{s}"]
            a = F.relu(F.conv2d(a, W.reshape(W.shape[0], 3, 3, -1).permute(0, 3, 1, 2), b, padding=1))
        if l < L - 1:
            skips.append(a)
            a = F.avg_pool2d(a, 2)
    for l in range(L - 2, -1, -1):
        W, b = P[f"up{l}"]
        a = F.conv_transpose2d(a, W.reshape(2, 2, W.shape[1], -1).permute(3, 2, 0, 1), b, stride=2)
        a = torch.cat([a, skips[l]], dim=1)
        for s in "ab":
            W, b = P[f"dec{l}{s}"]
            a = F.relu(F.conv2d(a, W.reshape(W.shape[0], 3, 3, -1).permute(0, 3, 1, 2), b, padding=1))
    W, b = P["head"]
    return F.conv2d(a, W[:, :, None, None], b)


@pytest.mark.parametrize("levels,base,size", [(3, 4, 8), (4, 2, 16)])
def test_network_gradient_matches_torch_autograd(levels, base, size):
    kw = dict(in_ch=4, out_ch=2, base=base, levels=levels)
    flat = U.init_params(7, **kw)
    flat[U.param_count(**kw) - 2:] = [0.3, -0.2]   # nonzero head bias
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 4, size, size))
    tgt = rng.standard_normal((2, 2, size, size))
    w = np.array([1.0, 0.5])
    net = U.UNet(**kw)
    y = net.forward(flat, x)
    g = net.backward(y, tgt, w)
    pt = torch.from_numpy(flat.copy()).requires_grad_()
    yt = _torch_net(pt, torch.from_numpy(x), kw)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-10, atol=1e-12)
    d = yt - torch.from_numpy(tgt)
    loss = sum(w[c] * (d[:, c] ** 2).mean() for c in range(2))
    loss.backward()
    assert abs(net.loss(y, tgt, w) - loss.item()) < 1e-12 * abs(loss.item())
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-13)


def test_gradient_matches_finite_differences():
    kw = dict(in_ch=4, out_ch=2, base=2, levels=3)
    flat = U.init_params(11, **kw)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 4, 8, 8))
    tgt = rng.standard_normal((1, 2, 8, 8))
    w = np.array([1.0, 2.0])
    net = U.UNet(**kw)
    g = net.backward(net.forward(flat, x), tgt, w)
    idx = rng.choice(len(flat), 60, replace=False)
    h = 1e-6
    for i in idx:
        p = flat.copy()
        p[i] += h
        lp = net.loss(net.forward(p, x), tgt, w)
        p[i] -= 2 * h
        lm = net.loss(net.forward(p, x), tgt, w)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 3.0e38, -1e-40, np.inf, -np.inf]])
    want = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(U.bf16_round(v), want)
    assert np.isnan(U.bf16_round(np.array([np.nan]))[0])


def test_param_count_of_the_surrogate():
    """The 4-level, base-32 network of four inputs and two outputs."""
    n = U.param_count(in_ch=4, out_ch=2, base=32, levels=4)
    assert n == sum(int(np.prod(U.weight_shape(k, ci, co))) + co for _, k, ci, co in U.layers())
    assert 1.8e6 < n < 2.0e6
