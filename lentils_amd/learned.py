"""Learned preconditioner for the batched source inversion.

A small convolutional network (BF16 operands on MFMA, FP32 accumulate; the
kernels are in csrc/pcnet.hip) corrects the lens-structured preconditioner
M = c L_s^T L_s + lambda Cc^T Cc of the source inversion: for a CG residual r
of one lens sample, with m = M^-1 r (a few inner Jacobi-PCG steps), the
search direction is

    z = m + y(r / rms(r), m / rms(m)) rms(m),

so the network sees the residual and M's answer and outputs a correction
towards A^-1 r.  Its last layer starts at zero (an untrained network is M
alone).  It is trained on the GPU with the solver's own operator on
(right-hand side, solution) pairs the solver generates: b = A x for random
Gaussian-blob sources x over random lens models, and the residuals r_k of
truncated solves of those b (the pair (r_k, x - x_k)), by the energy loss
z^T A z - 2 z^T r (the A-norm error of z up to a constant, which needs no
explicit solution).  Inside flexible CG (BatchedSourceInversion.solve(
precond='learned')) the guard takes m whenever the corrected direction is
not a descent direction, and a sample that stalls switches the network off.
The reference has no counterpart (it preconditions with a sparse LU,
lentils/tests/test_all.py:63-72).
"""
import ctypes as C

import numpy as np

from . import backend as B


class PCNet:
    """The network on an H x W source grid, for batches of up to `batch`."""

    def __init__(self, H, W, batch=64, params=None, seed=0):
        self.H, self.W, self.batch = int(H), int(W), int(batch)
        h = C.c_void_p()
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.size != self.param_count():
                raise ValueError(f'expected {self.param_count()} parameters, got {p.size}')
            B.check(B.lib().lt_pcnet_create(self.H, self.W, self.batch, B.ptr(p), 0, C.byref(h)))
        else:
            B.check(B.lib().lt_pcnet_create(self.H, self.W, self.batch, None, int(seed), C.byref(h)))
        self._handle = h

    @staticmethod
    def param_count():
        return int(B.lib().lt_pcnet_param_count())

    @property
    def params(self):
        p = np.zeros(self.param_count(), dtype=np.float32)
        B.check(B.lib().lt_pcnet_get_params(self._handle, B.ptr(p)))
        return p

    @params.setter
    def params(self, value):
        p = np.ascontiguousarray(value, dtype=np.float32)
        B.check(B.lib().lt_pcnet_set_params(self._handle, B.ptr(p)))

    def save(self, path):
        np.save(path, self.params)

    @classmethod
    def load(cls, path, H, W, batch=64):
        return cls(H, W, batch, params=np.load(path, allow_pickle=False))

    def forward(self, x0):
        x = np.ascontiguousarray(x0, dtype=np.float32)
        n = x.shape[0]
        y = np.zeros((n, self.H, self.W), dtype=np.float32)
        for k in range(0, n, self.batch):
            m = min(self.batch, n - k)
            B.check(B.lib().lt_pcnet_forward_host(self._handle, B.ptr(x[k:k + m]), m, B.ptr(y[k:k + m])))
        return y

    def train(self, inputs, targets, epochs=10, batch=None, lr=2e-3, lr_final=1e-4, seed=0):
        """Adam on the mean relative squared error; returns the loss per epoch."""
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        hist = np.zeros(epochs)
        B.check(B.lib().lt_pcnet_train(self._handle, B.ptr(x), B.ptr(t), x.shape[0], int(epochs),
                                       int(batch or self.batch), float(lr), float(lr_final), int(seed), B.ptr(hist)))
        return hist

    def evaluate(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        out = C.c_double(0)
        B.check(B.lib().lt_pcnet_eval(self._handle, B.ptr(x), B.ptr(t), x.shape[0], C.byref(out)))
        return out.value

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and B._lib is not None:
            B._lib.lt_pcnet_destroy(h)
            self._handle = None


# the last layer (16 -> 1 channels, 3 x 3 kernels: 144 weights and one bias)
# closes the parameter vector (csrc/pcnet.hip: layers in order, each its
# [co][ci][3][3] weights then its biases); it is linear, so scaling it scales
# the network's correction
LAST_LAYER = 16 * 9 + 1


def scale_correction(params, alpha):
    """The parameters with the network's correction y scaled by alpha (the last
    layer's weights and bias times alpha; alpha = 0 gives M alone)."""
    p = np.array(params, dtype=np.float32, copy=True)
    p[-LAST_LAYER:] *= np.float32(alpha)
    return p


def blob_sources(src_space, n, rng, nblobs=4, spread=0.3, width=(0.05, 0.2), amp=(0.5, 1.0)):
    """n random sources, each a sum of Gaussian blobs evaluated at the source
    grid's pixel positions (the workload's source family)."""
    p = src_space.points
    out = np.zeros((n,) + p.shape[:2])
    for k in range(n):
        for _ in range(nblobs):
            c = rng.uniform(-spread, spread, 2)
            s = rng.uniform(*width)
            a = rng.uniform(*amp)
            out[k] += a * np.exp(-np.sum((p - c) ** 2, axis=-1) / (2 * s * s))
    return out


def solver_residuals(inversion, sources, iterations=(0, 4, 16, 48), precond='lens'):
    """Training vectors from `inversion` (a BatchedSourceInversion): for each
    draw in `sources` (list of per-sample source lists) the right-hand sides
    b = A x and the true residuals b - A x_k of the solver stopped after each
    k in `iterations` (k = 0 is b itself).  Returns a list of flat arrays of
    all samples; the inversion's own right-hand sides are restored."""
    keep = inversion.rhs()
    out = []
    try:
        for xs in sources:
            inversion.set_rhs(inversion.apply(xs))
            out.extend(inversion.residuals(iterations, precond=precond))
    finally:
        inversion.set_rhs(keep)
    return out
