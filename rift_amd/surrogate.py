"""U-Net flood surrogate (include/rift_unet.h) and ensemble screening.

`UNet` wraps a rift_unet handle; its inputs and outputs live in device
memory allocated through the same library (kp.DeviceArray), so that solver
maps reach the training set without a host round trip.  The screening workflow (`screen`) predicts the maximum-depth and
arrival-time maps of many scenarios, decides each scenario on a threshold of
its predicted depth at an asset, and re-runs with the solver every scenario
whose prediction lies within a margin of the threshold, plus a held-out share
that measures the surrogate's error.  The reference has no learned component;
the labels are the solver's own `save_max` / `save_arrival` maps
(peak_flood_depth / flood_wave_arrival, grid.cu:930-966).
"""
import ctypes as C

import numpy as np

from . import kp
from .kp import _check, load


def _addr(x):
    """Device address of a DeviceArray (element 0) or an integer pointer."""
    return C.c_void_p(x.ptr() if isinstance(x, kp.DeviceArray) else int(x))


class UNet:
    """A rift_unet handle: `levels` levels of base << l channels."""

    def __init__(self, height=256, width=256, in_channels=4, out_channels=2, base=32, levels=4,
                 max_batch=16, seed=1, device=0):
        L = load()
        self.desc = kp.UNetDesc(height, width, in_channels, out_channels, base, levels, max_batch,
                                seed, device)
        h = C.c_void_p()
        _check(L.rift_unet_create(C.byref(self.desc), C.byref(h)), "rift_unet_create")
        self.h = h
        self.height, self.width = height, width
        self.in_channels, self.out_channels = in_channels, out_channels
        self.max_batch = max_batch
        n = C.c_long()
        _check(L.rift_unet_param_count(self.h, C.byref(n)), "rift_unet_param_count")
        self.param_count = n.value

    def get_params(self):
        p = np.empty(self.param_count, np.float32)
        _check(load().rift_unet_get_params(self.h, p.ctypes.data_as(kp._F)), "rift_unet_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count:
            raise kp.RiftError(f"set_params: expected {self.param_count} values")
        _check(load().rift_unet_set_params(self.h, p.ctypes.data_as(kp._F)), "rift_unet_set_params")

    def set_loss_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.out_channels:
            raise kp.RiftError("set_loss_weights: one weight per output channel")
        _check(load().rift_unet_set_loss_weights(self.h, w.ctypes.data_as(kp._F)), "rift_unet_set_loss_weights")

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(load().rift_unet_set_adam(self.h, beta1, beta2, eps), "rift_unet_set_adam")

    def _batch(self, batch):
        if not 1 <= batch <= self.max_batch:
            raise kp.RiftError(f"batch must lie in [1, {self.max_batch}]")
        return int(batch)

    def forward(self, x, batch, y):
        """y = net(x) on device arrays (DeviceArray or address) of
        [batch][in_channels][H][W] and [batch][out_channels][H][W] floats."""
        _check(load().rift_unet_forward(self.h, _addr(x), self._batch(batch), _addr(y)), "rift_unet_forward")

    def gradient(self, x, target, batch):
        g = np.empty(self.param_count, np.float32)
        loss = C.c_double()
        _check(load().rift_unet_gradient(self.h, _addr(x), _addr(target), self._batch(batch), C.byref(loss),
                                         g.ctypes.data_as(kp._F)), "rift_unet_gradient")
        return loss.value, g

    def train_step(self, x, target, batch, lr):
        loss = C.c_double()
        _check(load().rift_unet_train_step(self.h, _addr(x), _addr(target), self._batch(batch), float(lr),
                                           C.byref(loss)), "rift_unet_train_step")
        return loss.value

    # host-array conveniences (one transfer each way)
    def _io(self, batch):
        xs = (batch, self.in_channels, self.height, self.width)
        ys = (batch, self.out_channels, self.height, self.width)
        return xs, ys

    def forward_np(self, x):
        x = np.ascontiguousarray(x, np.float32)
        xs, ys = self._io(x.shape[0])
        if x.shape != xs:
            raise kp.RiftError(f"forward_np: expected shape {xs}")
        dx, dy = kp.DeviceArray(xs, self.desc.device), kp.DeviceArray(ys, self.desc.device)
        dx.upload(x)
        self.forward(dx, xs[0], dy)
        return dy.download()

    def gradient_np(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        xs, ys = self._io(x.shape[0])
        if x.shape != xs or t.shape != ys:
            raise kp.RiftError("gradient_np: bad shapes")
        dx, dt = kp.DeviceArray(xs, self.desc.device), kp.DeviceArray(ys, self.desc.device)
        dx.upload(x)
        dt.upload(t)
        return self.gradient(dx, dt, xs[0])

    def flops(self, batch):
        f, t = C.c_double(), C.c_double()
        _check(load().rift_unet_flops(self.h, int(batch), C.byref(f), C.byref(t)), "rift_unet_flops")
        return f.value, t.value

    def close(self):
        if getattr(self, "h", None):
            load().rift_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
