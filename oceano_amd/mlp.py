"""The closure network (include/oceano_mlp.h) seen from Python: BF16 matrix-core training
and inference kernels of liboceano_dg.so on torch device tensors.

No counterpart in the reference (oceano has no learned closure)."""
import ctypes as C

import numpy as np
import torch

from . import _capi
from ._capi import check, fptr, lib

IN, HID, OUT = 32, 64, 32


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class ClosureNetwork:
    """n_in -> 3 x 64 (GELU) -> n_out; FP32 master weights, BF16 products."""

    def __init__(self, n_in, n_out, seed=0):
        h = C.c_void_p()
        check(lib().oc_mlp_create(n_in, n_out, seed, C.byref(h)), "oc_mlp_create")
        self._h, self.n_in, self.n_out = h, n_in, n_out
        self.n_params = int(lib().oc_mlp_n_params())
        self.normalization = (np.zeros(n_in), np.ones(n_in), np.zeros(n_out), np.ones(n_out))
        self.output_range = (np.full(n_out, -np.inf), np.full(n_out, np.inf))
        self.output_gain = 1.0

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().oc_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        check(lib().oc_mlp_get_params(self._h, fptr(p)), "oc_mlp_get_params")
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError("expected %d parameters" % self.n_params)
        check(lib().oc_mlp_set_params(self._h, fptr(p)), "oc_mlp_set_params")

    def set_normalization(self, x_mean, x_std, y_shift, y_scale):
        arrs = [np.ascontiguousarray(v, np.float32) for v in (x_mean, x_std, y_shift, y_scale)]
        self.normalization = tuple(a.astype(np.float64) for a in arrs)
        self._push_normalization()

    def _push_normalization(self):
        # the output gain scales the de-normalised output (y * scale + shift) as a whole; a
        # zero gain switches the correction off where the network is attached
        # (OceanoOperator.set_closure), so the kernels keep the unscaled output then
        xm, xs, ys, yc = self.normalization
        g = self.output_gain if self.output_gain > 0 else 1.0
        arrs = [np.ascontiguousarray(v, np.float32) for v in (xm, xs, g * ys, g * yc)]
        check(lib().oc_mlp_set_normalization(self._h, *[fptr(a) for a in arrs]),
              "oc_mlp_set_normalization")

    def set_output_gain(self, gain):
        """multiply the network's (de-normalised) output by `gain` >= 0: the a-posteriori
        calibration of the correction (0 = no correction: an operator given this network
        runs the plain model)"""
        if not gain >= 0:
            raise ValueError("output gain must be >= 0")
        self.output_gain = float(gain)
        self._push_normalization()

    def set_output_range(self, lo, hi):
        """saturation of the normalised output at inference (the trained range)"""
        a = np.ascontiguousarray(lo, np.float32)
        b = np.ascontiguousarray(hi, np.float32)
        check(lib().oc_mlp_set_output_range(self._h, fptr(a), fptr(b)), "oc_mlp_set_output_range")
        self.output_range = (a.astype(np.float64), b.astype(np.float64))

    @staticmethod
    def _f32(t):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("device float32 contiguous tensors expected")
        return t

    def forward(self, x):
        x = self._f32(x)
        y = torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        check(lib().oc_mlp_forward(self._h, x.shape[0], _vp(x), _vp(y), _stream()),
              "oc_mlp_forward")
        return y

    def gradients(self, x, t):
        """(loss, gradient as a host float32 array); synchronises"""
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        check(lib().oc_mlp_gradients(self._h, x.shape[0], _vp(self._f32(x)), _vp(self._f32(t)),
                                     fptr(g), C.byref(loss), _stream()), "oc_mlp_gradients")
        return loss.value, g

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().oc_mlp_adam(self._h, lr, beta1, beta2, eps, _stream()), "oc_mlp_adam")

    def train_step(self, x, t, lr, want_loss=False):
        loss = C.c_double()
        check(lib().oc_mlp_train_step(self._h, x.shape[0], _vp(self._f32(x)),
                                      _vp(self._f32(t)), lr,
                                      C.byref(loss) if want_loss else None, _stream()),
              "oc_mlp_train_step")
        return loss.value if want_loss else None

    def save(self, path, **extra):
        np.savez(path, params=self.get_params(), n_in=self.n_in, n_out=self.n_out,
                 x_mean=self.normalization[0], x_std=self.normalization[1],
                 y_shift=self.normalization[2], y_scale=self.normalization[3],
                 y_lo=self.output_range[0], y_hi=self.output_range[1],
                 output_gain=self.output_gain, **extra)

    @classmethod
    def load(cls, path):
        d = np.load(path)  # numeric arrays only (allow_pickle stays False)
        net = cls(int(d["n_in"]), int(d["n_out"]))
        net.set_params(d["params"])
        net.set_normalization(d["x_mean"], d["x_std"], d["y_shift"], d["y_scale"])
        if "y_lo" in d:
            net.set_output_range(d["y_lo"], d["y_hi"])
        if "output_gain" in d:
            net.set_output_gain(float(d["output_gain"]))
        return net
