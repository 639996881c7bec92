"""unet_oracle.py -- FP64 numpy restatement of the U-Net surrogate.

TEST INFRASTRUCTURE ONLY.  This is the checker the HIP surrogate kernels
(include/rift_unet.h) are compared against; nothing in the product loads it.
Only tests/ and __graft_entry__.smoke() may use it.

The surrogate has no counterpart in the reference (RIFT has no learned
component); the network is defined by the product's own parameter layout,
which this file restates:

  * activations NHWC; levels l = 0..L-1 with c_l = base * 2**l channels;
  * encoder level l: conv3x3(+bias, ReLU) twice; 2x2 average pool between
    levels;
  * decoder level l = L-2..0: 2x2 stride-2 transposed conv (+bias, no
    activation) from level l+1, concatenation [up, skip_l], conv3x3(+ReLU)
    twice;
  * head: 1x1 conv (+bias) to out_channels, FP32 output, planar [B][C][H][W].

Canonical parameter layout (rift_unet_get_params): layer by layer in the
order of `layers()`, each weight then its bias.  conv3x3 weight W[co][t][ci]
with tap t = 3*ky + kx reading input pixel (i + ky - 1, j + kx - 1), zero
padding; transposed conv W[s][co][ci], s = 2*a + b writing output pixel
(2i + a, 2j + b); head W[co][ci].

`bf16=True` rounds what the device stores in bfloat16 (weights, the input,
every activation) to the nearest even bfloat16, so that a comparison with the
device isolates accumulation order.  Loss: sum over channels c of w_c times
the mean over pixels of (y_c - target_c)**2.

Pins (tests/test_unet_oracle.py): every layer against torch's CPU float64
conv2d / conv_transpose2d / avg_pool2d and autograd, and the analytic
gradients against central finite differences.
"""
import numpy as np


def bf16_round(x):
    """Round float64 values to the nearest bfloat16 (ties to even) and return
    them as float64.  Goes through float32 first, as the device stores FP32
    results before rounding."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def layers(in_ch=4, out_ch=2, base=32, levels=4):
    """Layer table: (name, kind, cin, cout) in parameter order."""
    out = []
    c = [base << l for l in range(levels)]
    prev = in_ch
    for l in range(levels):
        out.append((f"enc{l}a", "conv3", prev, c[l]))
        out.append((f"enc{l}b", "conv3", c[l], c[l]))
        prev = c[l]
    for l in range(levels - 2, -1, -1):
        out.append((f"up{l}", "up2", c[l + 1], c[l]))
        out.append((f"dec{l}a", "conv3", 2 * c[l], c[l]))
        out.append((f"dec{l}b", "conv3", c[l], c[l]))
    out.append(("head", "conv1", c[0], out_ch))
    return out


def weight_shape(kind, cin, cout):
    return {"conv3": (cout, 9, cin), "up2": (4, cout, cin), "conv1": (cout, cin)}[kind]


def param_count(**kw):
    n = 0
    for _, kind, cin, cout in layers(**kw):
        n += int(np.prod(weight_shape(kind, cin, cout))) + cout
    return n


def unpack(flat, **kw):
    """Flat canonical vector -> {name: (W, b)} (views)."""
    p, off = {}, 0
    for name, kind, cin, cout in layers(**kw):
        shp = weight_shape(kind, cin, cout)
        nw = int(np.prod(shp))
        p[name] = (flat[off:off + nw].reshape(shp), flat[off + nw:off + nw + cout])
        off += nw + cout
    assert off == len(flat)
    return p


def pack(grads, **kw):
    return np.concatenate([np.concatenate([grads[name][0].ravel(), grads[name][1].ravel()])
                           for name, _, _, _ in layers(**kw)])


def init_params(seed, **kw):
    """He-normal weights (fan-in), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, kind, cin, cout in layers(**kw):
        shp = weight_shape(kind, cin, cout)
        fan = {"conv3": 9 * cin, "up2": cin, "conv1": cin}[kind]
        parts.append(rng.standard_normal(int(np.prod(shp))) * np.sqrt(2.0 / fan))
        parts.append(np.zeros(cout))
    return np.concatenate(parts)


# --- layers (NHWC, float64) ------------------------------------------------

def conv3(x, W, b):
    B, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, Wd, W.shape[0]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        y += xp[:, ky:ky + H, kx:kx + Wd, :] @ W[:, t, :].T
    return y + b


def conv3_backward(x, W, dy):
    B, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    dyf = dy.reshape(-1, dy.shape[-1])
    for t in range(9):
        ky, kx = divmod(t, 3)
        patch = xp[:, ky:ky + H, kx:kx + Wd, :].reshape(-1, x.shape[-1])
        dW[:, t, :] = dyf.T @ patch
        dxp[:, ky:ky + H, kx:kx + Wd, :] += dy @ W[:, t, :]
    return dxp[:, 1:H + 1, 1:Wd + 1, :], dW, dyf.sum(0)


def up2(x, W, b):
    B, H, Wd, _ = x.shape
    y = np.zeros((B, 2 * H, 2 * Wd, W.shape[1]))
    for s in range(4):
        a, c = divmod(s, 2)
        y[:, a::2, c::2, :] = x @ W[s].T
    return y + b


def up2_backward(x, W, dy):
    dx = np.zeros_like(x)
    dW = np.zeros_like(W)
    xf = x.reshape(-1, x.shape[-1])
    for s in range(4):
        a, c = divmod(s, 2)
        g = dy[:, a::2, c::2, :]
        dx += g @ W[s]
        dW[s] = g.reshape(-1, g.shape[-1]).T @ xf
    return dx, dW, dy.reshape(-1, dy.shape[-1]).sum(0)


def pool2(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def pool2_backward(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


# --- network ---------------------------------------------------------------

class UNet:
    """Forward/backward of the surrogate in float64."""

    def __init__(self, in_ch=4, out_ch=2, base=32, levels=4):
        self.kw = dict(in_ch=in_ch, out_ch=out_ch, base=base, levels=levels)
        self.levels = levels

    def forward(self, flat, x_planar, bf16=False):
        """x_planar [B][Cin][H][W] -> y [B][Cout][H][W]; keeps the tape."""
        R = bf16_round if bf16 else (lambda v: v)
        # weights in bfloat16, biases FP32 (as the device keeps them)
        P = {k: (R(w), b) for k, (w, b) in unpack(np.asarray(flat, np.float64), **self.kw).items()}
        L = self.levels
        tape = {}
        a = R(np.transpose(np.asarray(x_planar, np.float64), (0, 2, 3, 1)))
        tape["x"] = a
        skips = []
        for l in range(L):
            for s in "ab":
                W, b = P[f"enc{l}{s}"]
                tape[f"enc{l}{s}.in"] = a
                a = R(np.maximum(conv3(a, W, b), 0.0))
                tape[f"enc{l}{s}"] = a
            if l < L - 1:
                skips.append(a)
                a = R(pool2(a))
                tape[f"pool{l}"] = a
        for l in range(L - 2, -1, -1):
            W, b = P[f"up{l}"]
            tape[f"up{l}.in"] = a
            u = R(up2(a, W, b))
            tape[f"up{l}"] = u
            a = np.concatenate([u, skips[l]], axis=-1)
            for s in "ab":
                W, b = P[f"dec{l}{s}"]
                tape[f"dec{l}{s}.in"] = a
                a = R(np.maximum(conv3(a, W, b), 0.0))
                tape[f"dec{l}{s}"] = a
        W, b = P["head"]
        tape["head.in"] = a
        y = a @ W.T + b
        self.tape, self.P, self.R = tape, P, R
        return np.transpose(y, (0, 3, 1, 2))

    def loss(self, y, target, weights):
        B, C, H, W = y.shape
        d = y - target
        return float(sum(weights[c] * np.mean(d[:, c] ** 2) for c in range(C)))

    def backward(self, y, target, weights):
        """Gradient of loss() w.r.t. the canonical parameters (after forward).
        With bf16 the device rounds the head gradient and every propagated
        activation gradient to bfloat16; so does this."""
        B, C, H, W = y.shape
        wv = np.asarray(weights, np.float64).reshape(1, C, 1, 1)
        g = self.R(np.transpose(2.0 * wv * (y - target) / (B * H * W), (0, 2, 3, 1)))
        R, P, T, L = self.R, self.P, self.tape, self.levels
        C = g.shape[-1]
        grads = {}
        Wh, _ = P["head"]
        hin = T["head.in"]
        grads["head"] = (g.reshape(-1, C).T @ hin.reshape(-1, hin.shape[-1]), g.reshape(-1, C).sum(0))
        d = R((g @ Wh) * (hin > 0))                     # d(dec0b)
        dskip = {}
        for l in range(0, L - 1):
            c = P[f"dec{l}b"][0].shape[0]
            W, _ = P[f"dec{l}b"]
            dx, dW, db = conv3_backward(T[f"dec{l}b.in"], W, d)
            grads[f"dec{l}b"] = (dW, db)
            d = R(dx * (T[f"dec{l}a"] > 0))             # d(dec{l}a)
            W, _ = P[f"dec{l}a"]
            dx, dW, db = conv3_backward(T[f"dec{l}a.in"], W, d)
            grads[f"dec{l}a"] = (dW, db)
            dup, dskip[l] = R(dx[..., :c]), R(dx[..., c:])
            W, _ = P[f"up{l}"]
            dx, dW, db = up2_backward(T[f"up{l}.in"], W, dup)
            grads[f"up{l}"] = (dW, db)
            src = f"dec{l + 1}b" if l + 1 < L - 1 else f"enc{L - 1}b"
            d = R(dx * (T[src] > 0))                    # d(dec{l+1}b) or d(enc{L-1}b)
        # d = gradient of the bottleneck output; walk the encoder down
        for l in range(L - 1, -1, -1):
            if l < L - 1:
                d = R((dskip[l] + pool2_backward(d)) * (T[f"enc{l}b"] > 0))
            W, _ = P[f"enc{l}b"]
            dx, dW, db = conv3_backward(T[f"enc{l}b.in"], W, d)
            grads[f"enc{l}b"] = (dW, db)
            d = R(dx * (T[f"enc{l}a"] > 0))
            W, _ = P[f"enc{l}a"]
            dx, dW, db = conv3_backward(T[f"enc{l}a.in"], W, d)
            grads[f"enc{l}a"] = (dW, db)
            d = R(dx)                                   # d(pool{l-1})
        return pack(grads, **self.kw)
