"""CPU oracle of the closure network (test infrastructure only: the tests use it as the
checker of the GPU kernels; the product never imports it).

Restates, in numpy, the arithmetic of oceano_amd/csrc/mlp_core.hpp and mlp.hip: inputs
normalised in FP32 and rounded to BF16 (round to nearest even), products of BF16 operands
accumulated in FP32 (here in FP64, then rounded: the order of the matrix-core accumulation
is not reproduced, so comparisons are to a tolerance), bias + GELU (erf form) in FP32 with
BF16 activations, MSE of the normalised outputs, and the backward pass with BF16 output /
pre-activation gradients as the products' operands and FP32 bias gradients.

`quantize=False` runs the same formulas in plain FP64: the finite-difference checks of
the tests pin the backward formulas in that mode.  The network has no counterpart in the
reference (oceano has no learned closure).
"""
import math

import numpy as np

IN, HID, OUT = 32, 64, 32
W1 = 0
B1 = W1 + IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * HID
W4 = B3 + HID
B4 = W4 + HID * OUT
NPARAM = B4 + OUT

_erf = np.vectorize(math.erf, otypes=[np.float64])


def to_bf16(x):
    """float32 -> nearest-even BF16, returned as float32 values"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    return [(p[W1:B1].reshape(IN, HID), p[B1:W2]), (p[W2:B2].reshape(HID, HID), p[B2:W3]),
            (p[W3:B3].reshape(HID, HID), p[B3:W4]), (p[W4:B4].reshape(HID, OUT), p[B4:])]


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(
        2.0 * math.pi)


class Network:
    def __init__(self, params, n_in, n_out, norm=None, quantize=True):
        self.p = np.asarray(params, np.float64)
        self.n_in, self.n_out, self.quantize = n_in, n_out, quantize
        if norm is None:
            norm = (np.zeros(n_in), np.ones(n_in), np.zeros(n_out), np.ones(n_out))
        self.x_mean, self.x_std, self.y_shift, self.y_scale = [np.asarray(v, np.float64)
                                                               for v in norm]

    def _q(self, a):
        return to_bf16(a).astype(np.float64) if self.quantize else a

    def _f(self, a):
        return np.asarray(a, np.float32).astype(np.float64) if self.quantize else a

    def _inputs(self, x):
        n = x.shape[0]
        X = np.zeros((n, IN))
        if self.quantize:
            inv = (1.0 / self.x_std.astype(np.float32)).astype(np.float32)
            xn = (x.astype(np.float32) - self.x_mean.astype(np.float32)) * inv
            X[:, :self.n_in] = xn.astype(np.float64)
        else:
            X[:, :self.n_in] = (x - self.x_mean) / self.x_std
        return self._q(X)

    def forward(self, x, keep=False):
        """normalised output [n, OUT] (and the cache for backward)"""
        layers = unpack(self.p)
        h = self._inputs(np.asarray(x, np.float64))
        cache = {"X": h, "Z": [], "H": [h]}
        for li, (W, b) in enumerate(layers):
            z = self._f(self._f(h @ self._q(self._f(W))) + self._f(b))
            if li < 3:
                cache["Z"].append(z)
                h = self._q(self._f(gelu(z)))
                cache["H"].append(h)
            else:
                y = z
        return (y, cache) if keep else y

    def predict(self, x, lo=None, hi=None):
        """de-normalised output; lo / hi saturate the normalised output (inference)"""
        y = self.forward(x)[:, :self.n_out]
        if lo is not None:
            y = np.clip(y, lo, hi)
        return y * self.y_scale + self.y_shift

    def loss_and_gradient(self, x, t):
        """MSE of the normalised outputs over n x n_out and its gradient (NPARAM)"""
        x = np.asarray(x, np.float64)
        n = x.shape[0]
        y, cache = self.forward(x, keep=True)
        tn = np.zeros((n, OUT))
        tn[:, :self.n_out] = (np.asarray(t, np.float64) - self.y_shift) / self.y_scale
        tn = self._f(tn)
        d = np.zeros((n, OUT))
        d[:, :self.n_out] = (y - tn)[:, :self.n_out]
        loss = float((d ** 2).sum() / (n * self.n_out))
        gscale = 2.0 / (n * self.n_out)
        if self.quantize:
            gscale = float(np.float32(2.0) / np.float32(float(n) * float(self.n_out)))
        g = self._f(d * gscale)
        layers = unpack(self.p)
        grad = np.zeros(NPARAM)
        offs = [(W1, B1), (W2, B2), (W3, B3), (W4, B4)]
        dz_f, dz_q = g, self._q(g)
        for li in range(3, -1, -1):
            W, _ = layers[li]
            Hin = cache["H"][li]
            wo, bo = offs[li]
            grad[bo:bo + W.shape[1]] = dz_f.sum(0)
            grad[wo:wo + W.size] = (Hin.T @ dz_q).ravel()
            if li == 0:
                break
            dh = self._f(dz_q @ self._q(self._f(W)).T)
            dz_f = self._f(dh * self._f(gelu_prime(cache["Z"][li - 1])))
            dz_q = self._q(dz_f)
        return loss, grad


def adam(p, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """one bias-corrected Adam step (FP32 state), as k_mlp_adam"""
    f = np.float32
    lr_t = f(lr * math.sqrt(1.0 - beta2 ** step) / (1.0 - beta1 ** step))
    m = f(beta1) * m.astype(f) + (f(1) - f(beta1)) * g.astype(f)
    v = f(beta2) * v.astype(f) + (f(1) - f(beta2)) * g.astype(f) * g.astype(f)
    p = p.astype(f) - lr_t * m / (np.sqrt(v) + f(eps))
    return p, m, v
