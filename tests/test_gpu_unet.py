"""The surrogate's HIP kernels through the C ABI against the FP64 oracle
(oracle/unet_oracle.py, itself pinned to torch CPU float64 and finite
differences): forward outputs, loss and every layer's parameter gradient.

The oracle runs with the device's bfloat16 rounding of weights, input,
activations and propagated gradients, so what is left is FP32 against FP64
accumulation: now and then it moves a rounded bfloat16 value by one unit
(0.4 %), and in a deep network such a change spreads through the layers after
it.  A two-level network shows the kernels agree to rounding: forward
relative L2 error 1e-5, every gradient tensor 2e-3 (measured 2e-8 in the
decoder, <= 4e-4 in the encoder).  The four-level network of the surrogate
(bottleneck 4 x 4 at 32 x 32) is held to 1e-2 forward and 6e-2 per gradient
tensor (measured <= 4.4e-2 at the bottleneck), 2.5e-2 on average.  So that
an error confined to one wide layer cannot hide in that spread, two-level
networks of 64 and 128 base channels run every kernel instantiation of the
deeper levels (N up to 256) layer by layer at 3e-3 .. 6e-3 forward and
2e-2 .. 4e-2 per gradient tensor.  A transposed operand, a wrong tap or a
wrong sub-pixel phase gives errors of order one.
"""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu

KW = dict(in_ch=4, out_ch=2, base=32, levels=4)


def _kw(levels, base=32):
    return dict(KW, levels=levels, base=base)


def _net(surr, size, batch, seed=3, levels=4, base=32):
    return surr.UNet(size, size, 4, 2, base, levels, max_batch=batch, seed=seed)


def _params(seed, levels=4, base=32):
    flat = U.init_params(seed, **_kw(levels, base))
    rng = np.random.default_rng(seed + 100)
    off = 0
    for _, kind, cin, cout in U.layers(**_kw(levels, base)):      # nonzero biases
        off += int(np.prod(U.weight_shape(kind, cin, cout)))
        flat[off:off + cout] = 0.1 * rng.standard_normal(cout)
        off += cout
    return flat.astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def surr(rift):
    from rift_amd import surrogate
    return surrogate


@pytest.mark.parametrize("levels,size,batch,tol", [(2, 16, 2, 1e-5), (4, 32, 2, 1e-2), (4, 48, 3, 1e-2)])
def test_forward_matches_oracle(surr, levels, size, batch, tol):
    print("forward", levels, size, batch, flush=True)
    net = _net(surr, size, batch, levels=levels)
    flat = _params(5, levels)
    net.set_params(flat)
    np.testing.assert_array_equal(net.get_params(), flat)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((batch, 4, size, size)).astype(np.float32)
    y = net.forward_np(x)
    yo = U.UNet(**_kw(levels)).forward(flat.astype(np.float64), x, bf16=True)
    print("rel", _rel(y, yo))
    assert _rel(y, yo) < tol, _rel(y, yo)
    assert np.abs(y - yo).max() < 3 * tol * (yo.max() - yo.min())
    # a second call gives the same result
    y2 = net.forward_np(x)
    np.testing.assert_array_equal(y, y2)
    net.close()


@pytest.mark.parametrize("levels,size,tol,mean_tol", [(2, 16, 2e-3, 1e-3), (4, 32, 6e-2, 2.5e-2)])
def test_gradient_matches_oracle(surr, levels, size, tol, mean_tol):
    batch = 2
    kw = _kw(levels)
    net = _net(surr, size, batch, levels=levels)
    flat = _params(6, levels)
    net.set_params(flat)
    net.set_loss_weights([1.0, 0.5])
    rng = np.random.default_rng(10)
    x = rng.standard_normal((batch, 4, size, size)).astype(np.float32)
    tgt = rng.standard_normal((batch, 2, size, size)).astype(np.float32)
    loss, g = net.gradient_np(x, tgt)
    orc = U.UNet(**kw)
    y = orc.forward(flat.astype(np.float64), x, bf16=True)
    lo = orc.loss(y, tgt, [1.0, 0.5])
    go = orc.backward(y, tgt, [1.0, 0.5])
    assert abs(loss - lo) < 1e-2 * lo, (loss, lo)
    gp, gop = U.unpack(g.astype(np.float64), **kw), U.unpack(go, **kw)
    bad, errs = [], []
    for name in gp:
        for part, a, b in (("W", gp[name][0], gop[name][0]), ("b", gp[name][1], gop[name][1])):
            e = _rel(a, b)
            errs.append(e)
            print(f"{name}.{part} rel {e:.2e}")
            if e > tol:
                bad.append((name, part, e))
    assert not bad, bad
    assert np.mean(errs) < mean_tol, np.mean(errs)
    net.close()


def test_training_reduces_the_loss(surr):
    """Adam on one fixed batch: the loss falls well below its start."""
    from rift_amd import kp
    size, batch = 32, 4
    net = _net(surr, size, batch, seed=11)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((batch, 4, size, size)).astype(np.float32)
    # a smooth target that depends on the input
    t = np.stack([x[:, 0] * 0.5 + x[:, 1] * 0.25, np.tanh(x[:, 2])], axis=1).astype(np.float32)
    dx, dt = kp.DeviceArray(x.shape), kp.DeviceArray(t.shape)
    dx.upload(x)
    dt.upload(t)
    losses = [net.train_step(dx, dt, batch, 2e-3) for _ in range(60)]
    print("loss", losses[0], losses[-1], flush=True)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.2 * losses[0]
    net.close()


def test_rejects_bad_arguments(surr):
    from rift_amd import kp
    net = _net(surr, 32, 2)
    with pytest.raises(kp.RiftError):
        net.forward_np(np.zeros((3, 4, 32, 32), np.float32))     # batch > max_batch
    with pytest.raises(kp.RiftError):
        net.forward_np(np.zeros((2, 4, 16, 32), np.float32))
    with pytest.raises(kp.RiftError):
        surr.UNet(30, 32, max_batch=1)                            # not a multiple of 8
    net.close()


@pytest.mark.parametrize("base,tol_y,tol_g,mean_tol", [(64, 3e-3, 2e-2, 5e-3), (128, 6e-3, 4e-2, 1e-2)])
def test_wide_two_level_nets_match_oracle_per_layer(surr, base, tol_y, tol_g, mean_tol):
    """Every kernel instantiation of the four-level surrogate, checked layer by
    layer in a network too shallow for rounding to pile up: a two-level
    network of 64 (128) base channels runs the N = 64 / 128 (128 / 256) 3 x 3
    kernels, their mirrored input-gradient forms, the transposed
    convolutions and the weight gradients at the widths of the surrogate's
    deeper levels.  With K = 9 x 64 .. 9 x 256 terms, FP32 sums differ from
    FP64 ones by ~1e-6 relative, which moves the bfloat16 rounding of a few
    activations by one unit (0.4 %), and that change spreads to its
    neighbours: measured forward relative L2 error 9.7e-4 (base 64) and
    2.3e-3 (128), every parameter tensor's gradient <= 6.4e-3 and <= 1.6e-2
    (the first layer's weights, after the longest chain of layers), 2e-3 and
    5e-3 on average; held to about twice that.  A transposed operand, a
    wrong tap or phase gives errors of order one."""
    levels, size, batch = 2, 16, 2
    kw = _kw(levels, base)
    net = _net(surr, size, batch, levels=levels, base=base)
    flat = _params(7, levels, base)
    net.set_params(flat)
    net.set_loss_weights([1.0, 0.5])
    rng = np.random.default_rng(11)
    x = rng.standard_normal((batch, 4, size, size)).astype(np.float32)
    tgt = rng.standard_normal((batch, 2, size, size)).astype(np.float32)
    orc = U.UNet(**kw)
    y = net.forward_np(x)
    yo = orc.forward(flat.astype(np.float64), x, bf16=True)
    print(f"base {base} forward rel {_rel(y, yo):.2e}", flush=True)
    assert _rel(y, yo) < tol_y, _rel(y, yo)
    loss, g = net.gradient_np(x, tgt)
    lo = orc.loss(yo, tgt, [1.0, 0.5])
    go = orc.backward(yo, tgt, [1.0, 0.5])
    assert abs(loss - lo) < 1e-3 * lo, (loss, lo)
    gp, gop = U.unpack(g.astype(np.float64), **kw), U.unpack(go, **kw)
    bad, errs = [], []
    for name in gp:
        for part, a, b in (("W", gp[name][0], gop[name][0]), ("b", gp[name][1], gop[name][1])):
            e = _rel(a, b)
            errs.append(e)
            print(f"{name}.{part} rel {e:.2e}")
            if e > tol_g:
                bad.append((name, part, e))
    assert not bad, bad
    assert np.mean(errs) < mean_tol, np.mean(errs)
    net.close()
