"""The per-tap Kaiser-Bessel polynomials the HIP gridder evaluates in place
of kb_g (lentils/backend/nufft.c:11-21) reproduce the oracle's kernel.

lt_kb_taps runs the library's host-side fit (no GPU call).  The check
evaluates the polynomials the way the kernels do (Horner in the working
precision) against the oracle's I0 power series, which is itself pinned to
scipy.special.i0 in test_oracle_nufft.py."""
import ctypes as C

import numpy as np
import pytest

from lentils_amd import backend as B



def _deg(prec, W):
    return 16 if prec == B.LT_F64 else (9 if W <= 2 else 7)


def _taps(beta, W, prec):
    D = _deg(prec, W)
    c = np.zeros((2 * W + 1) * (D + 1))
    err = C.c_double(0.0)
    B.check(B.lib().lt_kb_taps(float(beta), int(W), int(prec), B.ptr(c), C.byref(err)))
    return c.reshape(2 * W + 1, D + 1), err.value


def _horner(c, f, dtype):
    f = f.astype(dtype)
    acc = np.full_like(f, dtype(c[-1]))
    for m in range(len(c) - 2, -1, -1):
        acc = acc * f + dtype(c[m])
    return acc.astype(np.float64)


@pytest.mark.parametrize('W,beta', [(3, 14.04), (4, None), (2, None), (6, None), (8, None), (1, 2.0)])
@pytest.mark.parametrize('prec', [B.LT_F32, B.LT_F64])
def test_taps_match_oracle_kernel(oracle, W, beta, prec):
    if beta is None:
        beta = oracle.kb_beta_default(W, 2)
    c, err = _taps(beta, W, prec)
    tol = 5e-7 if prec == B.LT_F32 else 1e-13
    dtype = np.float32 if prec == B.LT_F32 else np.float64
    peak = oracle.bessel_i0(beta)
    f = np.linspace(-0.5, 0.5, 401)
    worst = 0.0
    for k in range(2 * W + 1):
        x = f + W - k
        got = _horner(c[k], f, dtype)
        if k == 0:
            got[f > 0] = 0.0
        if k == 2 * W:
            got[f < 0] = 0.0
        want = np.array([oracle.bessel_i0(beta * np.sqrt(max(0.0, 1.0 - (xi / W) ** 2))) if abs(xi) <= W else 0.0
                         for xi in x])
        worst = max(worst, np.max(np.abs(got - want)) / peak)
    assert worst <= tol, worst
    assert err <= (1e-6 if prec == B.LT_F32 else 1e-13)


def test_edge_taps_at_zero_offset(oracle):
    """f = 0: both edge taps sit exactly at |x| = W, where the reference's
    kernel is I0(0) = 1 (|k| <= W is inclusive, nufft.c:12)."""
    peak = oracle.bessel_i0(14.04)
    for prec, tol in ((B.LT_F32, 5e-7), (B.LT_F64, 1e-13)):
        c, _ = _taps(14.04, 3, prec)
        assert abs(c[0][0] - 1.0) / peak < tol and abs(c[6][0] - 1.0) / peak < tol
        assert abs(c[3][0] - peak) / peak < tol


def test_bad_arguments_rejected():
    err = C.c_double(0.0)
    c = np.zeros(1000)
    assert B.lib().lt_kb_taps(14.04, 0, B.LT_F32, B.ptr(c), C.byref(err)) == B.LT_ERR_ARG
    assert B.lib().lt_kb_taps(-1.0, 3, B.LT_F32, B.ptr(c), C.byref(err)) == B.LT_ERR_ARG


def _window_taps(beta, W, prec):
    D = _deg(prec, W)
    c = np.zeros(2 * W * (D + 1))
    err = C.c_double(0.0)
    B.check(B.lib().lt_kb_taps_window(float(beta), int(W), int(prec), B.ptr(c), C.byref(err)))
    return c.reshape(2 * W, D + 1), err.value


def _even_odd(c, s, dtype):
    """The kernels' evaluation: E(s^2) +- s O(s^2) in the working precision."""
    s = s.astype(dtype)
    q = s * s
    D = len(c) - 1
    DE, DO = D // 2 * 2, (D if D % 2 else D - 1)
    e = np.full_like(s, dtype(c[DE]))
    for m in range(DE - 2, -1, -2):
        e = e * q + dtype(c[m])
    o = np.full_like(s, dtype(c[DO]))
    for m in range(DO - 2, 0, -2):
        o = o * q + dtype(c[m])
    return (e + s * o).astype(np.float64), (e - s * o).astype(np.float64)


@pytest.mark.parametrize('W,beta', [(3, 14.04), (4, None), (2, None), (6, None), (8, None), (1, 2.0)])
@pytest.mark.parametrize('prec', [B.LT_F32, B.LT_F64])
def test_window_taps_match_oracle_kernel(oracle, W, beta, prec):
    """The degridder's window taps: tap j of the 2W cells from floor(x)-W+1
    at s = x - floor(x) - 1/2 equals the oracle's kernel at distance
    s + W - 1/2 - j, with taps 2W-1-j evaluated as tap j at -s (the kernel is
    even) exactly as the kernels do; a distance-to-tap mix-up (a tap shifted
    by one cell or mirrored the wrong way) misses by far more than the bar."""
    if beta is None:
        beta = oracle.kb_beta_default(W, 2)
    c, err = _window_taps(beta, W, prec)
    tol = 5e-7 if prec == B.LT_F32 else 1e-13
    dtype = np.float32 if prec == B.LT_F32 else np.float64
    peak = oracle.bessel_i0(beta)
    s = np.linspace(-0.5, 0.5, 401)
    worst = 0.0
    for j in range(W):
        plus, minus = _even_odd(c[j], s, dtype)
        for tap, got in ((j, plus), (2 * W - 1 - j, minus)):
            x = s + W - 0.5 - tap
            want = np.array([oracle.bessel_i0(beta * np.sqrt(max(0.0, 1.0 - (xi / W) ** 2))) for xi in x])
            worst = max(worst, np.max(np.abs(got - want)) / peak)
    assert worst <= tol, worst
    assert err <= (1e-6 if prec == B.LT_F32 else 1e-13)
    # the stored fits of the upper taps are the mirrored lower ones
    for j in range(W):
        a = np.polynomial.polynomial.polyval(s, c[2 * W - 1 - j])
        b = np.polynomial.polynomial.polyval(-s, c[j])
        assert np.max(np.abs(a - b)) / peak <= 2 * tol
