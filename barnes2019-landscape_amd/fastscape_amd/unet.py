"""Binding of the learned proposal network (fs_unet_* in include/fastscape.h).

A 4-level convolutional U-Net in BF16 on the MFMA units with FP32
accumulation; forward, backward and Adam are the library's own HIP kernels.
Inputs per sample: an H x W initial terrain and 3 normalised parameters;
output: one H x W channel.
"""
import ctypes

import numpy as np

from ._lib import _check, _lib, _ptr

_f = ctypes.c_float


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class UNet:
    def __init__(self, height, width, max_batch=16, seed=0, device=0):
        self._u = None
        self.H, self.W, self.max_batch = int(height), int(width), int(max_batch)
        u = ctypes.c_void_p()
        _check(_lib().fs_unet_create(self.H, self.W, self.max_batch, int(seed), int(device),
                                     ctypes.byref(u)), "fs_unet_create")
        self._u = u
        n = ctypes.c_size_t(0)
        _check(_lib().fs_unet_param_count(u, ctypes.byref(n)), "fs_unet_param_count")
        self.nparam = n.value

    def close(self):
        if self._u is not None:
            _lib().fs_unet_destroy(self._u)
            self._u = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layers(self):
        """[(cin, cout, level)] of the 14 conv layers, in parameter order."""
        n = ctypes.c_int(0)
        _check(_lib().fs_unet_layer_shapes(self._u, None, None, None, 0, ctypes.byref(n)),
               "fs_unet_layer_shapes")
        a = np.zeros((3, n.value), dtype=np.int32)
        _check(_lib().fs_unet_layer_shapes(self._u, _ptr(a[0], ctypes.c_int32),
                                           _ptr(a[1], ctypes.c_int32), _ptr(a[2], ctypes.c_int32),
                                           n.value, ctypes.byref(n)), "fs_unet_layer_shapes")
        return [tuple(int(v) for v in a[:, i]) for i in range(n.value)]

    def get_params(self):
        p = np.empty(self.nparam, dtype=np.float32)
        _check(_lib().fs_unet_get_params(self._u, _ptr(p, _f), p.size), "fs_unet_get_params")
        return p

    def set_params(self, p):
        p = _f32(p).ravel()
        _check(_lib().fs_unet_set_params(self._u, _ptr(p, _f), p.size), "fs_unet_set_params")

    def forward(self, terrain, params):
        """terrain (B, H, W), params (B, 3) -> (B, H, W) float32."""
        terrain, params = _f32(terrain), _f32(params)
        B = terrain.shape[0]
        y = np.empty((B, self.H, self.W), dtype=np.float32)
        _check(_lib().fs_unet_forward(self._u, _ptr(terrain, _f), _ptr(params, _f), B,
                                      _ptr(y, _f)), "fs_unet_forward")
        return y

    def loss_grad(self, terrain, params, target):
        terrain, params, target = _f32(terrain), _f32(params), _f32(target)
        loss = _f(0)
        g = np.empty(self.nparam, dtype=np.float32)
        _check(_lib().fs_unet_loss_grad(self._u, _ptr(terrain, _f), _ptr(params, _f),
                                        _ptr(target, _f), terrain.shape[0], ctypes.byref(loss),
                                        _ptr(g, _f), g.size), "fs_unet_loss_grad")
        return float(loss.value), g

    def set_data(self, terrain, params, target):
        terrain, params, target = _f32(terrain), _f32(params), _f32(target)
        _check(_lib().fs_unet_set_data(self._u, _ptr(terrain, _f), _ptr(params, _f),
                                       _ptr(target, _f), terrain.shape[0]), "fs_unet_set_data")

    def train(self, epochs, batch, lr=1e-3, seed=0):
        losses = np.zeros(int(epochs), dtype=np.float32)
        _check(_lib().fs_unet_train(self._u, int(epochs), int(batch), float(lr), int(seed),
                                    _ptr(losses, _f)), "fs_unet_train")
        return losses

    def rank(self, terrain, params, target):
        """RMS interior misfit of each params row's output (one shared terrain)
        against target; outputs stay on the device."""
        terrain, params, target = _f32(terrain), _f32(params), _f32(target)
        mis = np.empty(params.shape[0], dtype=np.float32)
        _check(_lib().fs_unet_rank(self._u, _ptr(terrain, _f), _ptr(params, _f), params.shape[0],
                                   _ptr(target, _f), _ptr(mis, _f)), "fs_unet_rank")
        return mis

    def stats(self):
        ms, fl = ctypes.c_double(0), ctypes.c_double(0)
        n = ctypes.c_uint64(0)
        _check(_lib().fs_unet_stats(self._u, ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(n)),
               "fs_unet_stats")
        return dict(train_ms=ms.value, flops_per_sample=fl.value, samples=n.value)
