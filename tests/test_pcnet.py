"""The learned preconditioner network: the numpy oracle's backward pass is
pinned by finite differences and adjoint identities (CPU); the GPU forward
pass and the gradients of the hand-written BF16 MFMA backward kernels match
the oracle (GPU); training on the GPU lowers the loss (GPU)."""
import numpy as np
import pytest

from oracle import pcnet_ref as R


def _rng(seed):
    return np.random.default_rng(seed)


def test_oracle_adjoint_identities():
    g = _rng(1)
    x = g.normal(size=(2, 3, 8, 8))
    y = g.normal(size=(2, 3, 4, 4))
    assert np.isclose(np.sum(R._pool(x) * y), np.sum(x * R._pool_bwd(y)), rtol=1e-13)
    z = g.normal(size=(2, 3, 16, 16))
    assert np.isclose(np.sum(R._up(y) * x), np.sum(y * R._up_bwd(x)), rtol=1e-13)
    w = g.normal(size=(5, 3, 3, 3))
    dz = g.normal(size=(2, 5, 16, 16))
    lhs = np.sum(R._conv(z, w, None, False) * dz)
    rhs = np.sum(z * R._conv_dx(dz, w, False))
    assert np.isclose(lhs, rhs, rtol=1e-12)
    # the weight gradient is the derivative of <conv(x, w), dz> in w
    dw, db = R._conv_dw(dz, z, False)
    v = g.normal(size=w.shape)
    assert np.isclose(np.sum(R._conv(z, v, None, False) * dz), np.sum(dw * v), rtol=1e-12)
    assert np.allclose(db, dz.sum(axis=(0, 2, 3)))


def test_oracle_backward_matches_finite_differences():
    g = _rng(2)
    p = R.init_params(3)
    x0 = g.normal(size=(2, 2, 8, 8))
    dy = g.normal(size=(2, 8, 8))

    def f(q):
        return np.sum(R.forward(q, x0, bf16_ops=False)[0] * dy)

    y, cache = R.forward(p, x0, bf16_ops=False)
    grad = R.backward(p, cache, dy, bf16_ops=False)
    assert grad.shape == (R.param_count(),)
    eps = 1e-6
    for k in range(4):  # random directions
        v = g.normal(size=p.size)
        fd = (f(p + eps * v) - f(p - eps * v)) / (2 * eps)
        assert np.isclose(fd, grad @ v, rtol=1e-5, atol=1e-9), (fd, grad @ v)
    # single components of every layer (first weight, first bias)
    off = 0
    for ci, co in zip(R.CIN, R.COUT):
        for idx in (off, off + co * ci * 9):
            e = np.zeros(p.size)
            e[idx] = 1.0
            fd = (f(p + eps * e) - f(p - eps * e)) / (2 * eps)
            assert np.isclose(fd, grad[idx], rtol=1e-4, atol=1e-8), (idx, fd, grad[idx])
        off += co * ci * 9 + co


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-30, 65504.0], dtype=np.float64)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-30) / 3.0e-30 < 2 ** -8


def test_scale_correction_scales_the_output():
    """scale_correction scales exactly the last layer (the network's output is
    linear in it): in the oracle's network, alpha times the output."""
    from lentils_amd.learned import LAST_LAYER, scale_correction
    ci, co = R.CIN[-1], R.COUT[-1]
    assert LAST_LAYER == co * ci * 9 + co
    p = R.init_params(7)
    p[-LAST_LAYER:] = _rng(8).normal(size=LAST_LAYER) * 0.1  # a trained-looking last layer
    x0 = _rng(9).normal(size=(2, 2, 8, 8))
    y1, _ = R.forward(p, x0)
    for alpha in (0.0, 0.25, 1.0):
        q = scale_correction(p, alpha).astype(np.float64)
        assert np.array_equal(q[:-LAST_LAYER], p[:-LAST_LAYER].astype(np.float32).astype(np.float64))
        ya, _ = R.forward(q, x0)
        assert np.max(np.abs(ya - alpha * y1)) <= 1e-6 * np.max(np.abs(y1))


def _net(B, H, W, nb, params):
    import ctypes as C
    h = C.c_void_p()
    p32 = np.ascontiguousarray(params, dtype=np.float32)
    B.check(B.lib().lt_pcnet_create(H, W, nb, B.ptr(p32), 0, C.byref(h)))
    return h


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(hip):
    B = hip
    assert B.lib().lt_pcnet_param_count() == R.param_count()
    p = R.init_params(5).astype(np.float32).astype(np.float64)
    x0 = _rng(6).normal(size=(3, 2, 16, 24)).astype(np.float32)
    h = _net(B, 16, 24, 4, p)
    y = np.zeros((3, 16, 24), dtype=np.float32)
    B.check(B.lib().lt_pcnet_forward_host(h, B.ptr(x0), 3, B.ptr(y)))
    ref, _ = R.forward(p, x0.astype(np.float64), bf16_ops=True)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f'forward max error / max |y|: {err:.2e}')
    assert err < 2e-3
    B.lib().lt_pcnet_destroy(h)


@pytest.mark.gpu
def test_gpu_gradients_match_oracle(hip):
    B = hip
    p = R.init_params(7).astype(np.float32).astype(np.float64)
    g = _rng(8)
    x0 = g.normal(size=(2, 2, 16, 16)).astype(np.float32)
    dy = g.normal(size=(2, 16, 16)).astype(np.float32)
    h = _net(B, 16, 16, 2, p)
    y = np.zeros((2, 16, 16), dtype=np.float32)
    grad = np.zeros(R.param_count(), dtype=np.float32)
    B.check(B.lib().lt_pcnet_grad_host(h, B.ptr(x0), B.ptr(dy), 2, B.ptr(y), B.ptr(grad)))
    _, cache = R.forward(p, x0.astype(np.float64), bf16_ops=True)
    ref = R.backward(p, cache, dy.astype(np.float64), bf16_ops=True)
    off = 0
    for k, (ci, co) in enumerate(zip(R.CIN, R.COUT)):
        n = co * ci * 9 + co
        a, b = grad[off:off + n], ref[off:off + n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f'layer {k}: gradient relative error {rel:.2e}')
        assert rel < 2e-2, (k, rel)
        off += n
    B.lib().lt_pcnet_destroy(h)


@pytest.mark.gpu
def test_gpu_training_lowers_loss(hip):
    """A smoothing map (targets = 3x3 box blur of channel 0) is learnt by the
    GPU training loop: the held-out loss drops well below its initial value."""
    import ctypes as C
    B = hip
    g = _rng(9)
    n, H = 256, 16
    x = g.normal(size=(n, 2, H, H)).astype(np.float32)
    xp = np.pad(x[:, 0], ((0, 0), (1, 1), (1, 1)))
    t = sum(xp[:, a:a + H, b:b + H] for a in range(3) for b in range(3)).astype(np.float32) / 9.0
    h = C.c_void_p()
    B.check(B.lib().lt_pcnet_create(H, H, 32, None, 11, C.byref(h)))
    l0 = C.c_double(0)
    B.check(B.lib().lt_pcnet_eval(h, B.ptr(x[:64]), B.ptr(t[:64]), 64, C.byref(l0)))
    hist = np.zeros(8)
    B.check(B.lib().lt_pcnet_train(h, B.ptr(x[64:]), B.ptr(t[64:]), n - 64, 8, 32, 3e-3, 3e-4, 1, B.ptr(hist)))
    l1 = C.c_double(0)
    B.check(B.lib().lt_pcnet_eval(h, B.ptr(x[:64]), B.ptr(t[:64]), 64, C.byref(l1)))
    print(f'held-out loss {l0.value:.3f} -> {l1.value:.4f}; training history {np.round(hist, 4)}')
    assert l1.value < 0.2 * l0.value
    B.lib().lt_pcnet_destroy(h)
