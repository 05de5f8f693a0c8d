"""Test infrastructure only: CPU (numpy, FP64) restatement of the learned
preconditioner network of lentils_amd/csrc/pcnet.hip, used as the checker
of its forward pass and of the gradients its backward kernels produce.  The
network has no counterpart in the reference (which preconditions its CG with
a sparse LU, lentils/tests/test_all.py:63-72).

With ``bf16=True`` every convolution rounds its two operands to bfloat16
(round to nearest even) exactly where the GPU kernels do, so the GPU must
agree to FP32-accumulation rounding; with ``bf16=False`` the arithmetic is
plain FP64, which the finite-difference check of ``backward`` uses.

Parameter layout: per layer, weights [co][ci][3][3] then biases [co].
"""
import numpy as np

CIN = (2, 16, 32, 64, 128, 64, 32, 16)
COUT = (16, 32, 64, 128, 64, 32, 16, 1)


def param_count():
    return sum(co * ci * 9 + co for ci, co in zip(CIN, COUT))


def unpack(params):
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        w = params[off:off + co * ci * 9].reshape(co, ci, 3, 3)
        off += co * ci * 9
        b = params[off:off + co]
        off += co
        out.append((w, b))
    return out


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _conv(x, w, b, rnd):
    """3x3 convolution with zero padding: x [n][ci][H][W], w [co][ci][3][3]."""
    if rnd:
        x, w = bf16(x), bf16(w)
    n, ci, H, W = x.shape
    xp = np.zeros((n, ci, H + 2, W + 2))
    xp[:, :, 1:H + 1, 1:W + 1] = x
    y = np.zeros((n, w.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum('oc,ncij->noij', w[:, :, ky, kx], xp[:, :, ky:ky + H, kx:kx + W])
    if b is not None:
        y += b[None, :, None, None]
    return y


def _conv_dx(dz, w, rnd):
    """data gradient of _conv: correlation with the flipped, transposed kernel"""
    wf = np.ascontiguousarray(np.transpose(w[:, :, ::-1, ::-1], (1, 0, 2, 3)))
    return _conv(dz, wf, None, rnd)


def _conv_dw(dz, x, rnd):
    if rnd:
        dz, x = bf16(dz), bf16(x)
    n, ci, H, W = x.shape
    xp = np.zeros((n, ci, H + 2, W + 2))
    xp[:, :, 1:H + 1, 1:W + 1] = x
    dw = np.zeros((dz.shape[1], ci, 3, 3))
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = np.einsum('noij,ncij->oc', dz, xp[:, :, ky:ky + H, kx:kx + W])
    return dw, dz.sum(axis=(0, 2, 3))


def _pool(x):
    n, c, H, W = x.shape
    return 0.25 * (x[:, :, 0::2, 0::2] + x[:, :, 1::2, 0::2] + x[:, :, 0::2, 1::2] + x[:, :, 1::2, 1::2])


def _pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=2), 2, axis=3)


def _up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def _up_bwd(dy):
    return dy[:, :, 0::2, 0::2] + dy[:, :, 1::2, 0::2] + dy[:, :, 0::2, 1::2] + dy[:, :, 1::2, 1::2]


def _relu(z):
    return np.maximum(z, 0.0)


def forward(params, x0, bf16_ops=True):
    """x0: [n][2][H][W].  Returns (y [n][H][W], cache)."""
    L = unpack(np.asarray(params, dtype=np.float64))
    r = bf16_ops
    c = {'x0': np.asarray(x0, dtype=np.float64)}
    c['e1'] = _relu(_conv(c['x0'], *L[0], r))
    c['p1'] = _pool(c['e1'])
    c['e2'] = _relu(_conv(c['p1'], *L[1], r))
    c['p2'] = _pool(c['e2'])
    c['e3'] = _relu(_conv(c['p2'], *L[2], r))
    c['p3'] = _pool(c['e3'])
    c['e4'] = _relu(_conv(c['p3'], *L[3], r))
    c['u4'] = _up(c['e4'])
    c['a5'] = _relu(_conv(c['u4'], *L[4], r))
    c['d3'] = c['a5'] + c['e3']
    c['u3'] = _up(c['d3'])
    c['a6'] = _relu(_conv(c['u3'], *L[5], r))
    c['d2'] = c['a6'] + c['e2']
    c['u2'] = _up(c['d2'])
    c['a7'] = _relu(_conv(c['u2'], *L[6], r))
    c['d1'] = c['a7'] + c['e1']
    y = _conv(c['d1'], *L[7], r)[:, 0]
    return y, c


def backward(params, cache, dy, bf16_ops=True):
    """Gradient of sum(dy * y) w.r.t. the parameters (same layout)."""
    L = unpack(np.asarray(params, dtype=np.float64))
    r = bf16_ops
    c = cache
    g = [None] * 8
    dz = np.asarray(dy, dtype=np.float64)[:, None]
    g[7] = _conv_dw(dz, c['d1'], r)
    dd1 = _conv_dx(dz, L[7][0], r)
    de1 = dd1.copy()
    dz = dd1 * (c['a7'] > 0)
    g[6] = _conv_dw(dz, c['u2'], r)
    dd2 = _up_bwd(_conv_dx(dz, L[6][0], r))
    de2 = dd2.copy()
    dz = dd2 * (c['a6'] > 0)
    g[5] = _conv_dw(dz, c['u3'], r)
    dd3 = _up_bwd(_conv_dx(dz, L[5][0], r))
    de3 = dd3.copy()
    dz = dd3 * (c['a5'] > 0)
    g[4] = _conv_dw(dz, c['u4'], r)
    de4 = _up_bwd(_conv_dx(dz, L[4][0], r))
    dz = de4 * (c['e4'] > 0)
    g[3] = _conv_dw(dz, c['p3'], r)
    de3 = de3 + _pool_bwd(_conv_dx(dz, L[3][0], r))
    dz = de3 * (c['e3'] > 0)
    g[2] = _conv_dw(dz, c['p2'], r)
    de2 = de2 + _pool_bwd(_conv_dx(dz, L[2][0], r))
    dz = de2 * (c['e2'] > 0)
    g[1] = _conv_dw(dz, c['p1'], r)
    de1 = de1 + _pool_bwd(_conv_dx(dz, L[1][0], r))
    dz = de1 * (c['e1'] > 0)
    g[0] = _conv_dw(dz, c['x0'], r)
    return np.concatenate([np.concatenate([dw.reshape(-1), db]) for dw, db in g])


def init_params(seed=0):
    """He-initialised parameters (for tests; the library has its own)."""
    g = np.random.default_rng(seed)
    out = []
    for k, (ci, co) in enumerate(zip(CIN, COUT)):
        sd = np.sqrt(2.0 / (9 * ci)) * (0.1 if k == 7 else 1.0)
        out.append(g.normal(0, sd, size=co * ci * 9))
        out.append(g.normal(0, 0.05, size=co))
    return np.concatenate(out)
