"""GPU parity of the NUFFT (C ABI and the NUFFTOperator drop-in).

Bars: FP64 plan within 1e-12 (max-relative, test_utils.py:13-20) of the FP64
oracle; FP32 plan within 1e-5; both within the reference's own 1e-6 of the
brute-force DFT at the reference test's sizes (nufft_tests.py:43-67);
adjointness; run-to-run bitwise determinism of the atomic-free gridder."""
import ctypes as C

import numpy as np
import pytest

from helpers import annulus_uv, reference_test_uv, rng

pytestmark = pytest.mark.gpu


def _plan(B, uv, ch, nx, ny, bounds, wsup, beta, prec):
    from lentils_amd import ImageSpace, VisibilitySpace
    vs = VisibilitySpace(channels=ch, uvcoords=uv)
    ims = ImageSpace(shape=(nx, ny), bounds=bounds)
    h = C.c_void_p()
    B.check(B.lib().lt_nufft_create(vs, ims, 2, wsup, beta or 0.0, prec, C.byref(h)))
    return h, (vs, ims)


@pytest.mark.parametrize('wsup,beta,prec,tol', [(4, None, 0, 1e-12), (4, None, 1, 1e-5), (3, 14.04, 0, 1e-12),
                                                (3, 14.04, 1, 1e-5)])
def test_nufft_vs_oracle_reference_test_sizes(hip, oracle, wsup, beta, prec, tol):
    B, O = hip, oracle
    uv = reference_test_uv()
    h, keep = _plan(B, uv, [0.5], 236, 257, [(-0.5, 0.7), (-0.7, 0.5)], wsup, beta, prec)
    geo = O.NUFFTGeometry(O.Grid2D(236, 257, [-0.5, 0.7, -0.7, 0.5]), pad=2, wsup=wsup, beta=beta)
    g = rng(61)
    v = g.normal(size=(uv.shape[0], 2)).view(np.complex128).reshape(-1)
    img = np.zeros(236 * 257)
    B.check(B.lib().lt_nufft_adjoint_host(h, B.ptr(v), B.ptr(img)))
    ref = O.nufft_adjoint(v.reshape(1, -1), uv, [0.5], geo).reshape(-1)
    assert O.max_relative_error(img, ref) < tol
    x = g.normal(size=236 * 257)
    vis = np.zeros(uv.shape[0], dtype=np.complex128)
    B.check(B.lib().lt_nufft_forward_host(h, B.ptr(x), B.ptr(vis)))
    refv = O.nufft_forward(x.reshape(236, 257), uv, [0.5], geo).reshape(-1)
    assert O.max_relative_error(vis, refv) < tol
    B.lib().lt_nufft_destroy(h)


def test_nufft_operator_vs_dft(hip, oracle):
    """The reference's own test, through the drop-in operators."""
    from lentils_amd import DFTOperator, ImageSpace, NUFFTOperator, VisibilitySpace
    O = oracle
    uv = reference_test_uv()
    vs = VisibilitySpace(channels=[0.5], uvcoords=uv)
    ims = ImageSpace(shape=(236, 257), bounds=[(-0.5, 0.7), (-0.7, 0.5)])
    v = rng(62).normal(size=(uv.shape[0], 2)).view(np.complex128)
    nufft = NUFFTOperator(vs, ims)
    dft = DFTOperator(vs, ims)
    nufftvec = nufft.T * v
    dftvec = dft.T * v
    assert O.max_relative_error(nufftvec, dftvec) < 1e-6
    assert O.max_relative_error(nufft * dftvec, dft * dftvec) < 1e-6


@pytest.mark.parametrize('prec,tol', [(0, 1e-12), (1, 1e-5)])
def test_grid_degrid_c1_sizes(hip, oracle, prec, tol):
    B, O = hip, oracle
    uv = annulus_uv(100000)
    h, keep = _plan(B, uv, [1.0], 128, 128, [(-1.28, 1.28), (-1.28, 1.28)], 3, 14.04, prec)
    geo = O.NUFFTGeometry(O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28]), pad=2, wsup=3, beta=14.04)
    g = rng(63)
    grid = np.ascontiguousarray(g.normal(size=(256, 129, 2)).view(np.complex128)[..., 0])
    vis = np.zeros(100000, dtype=np.complex128)
    B.check(B.lib().lt_nufft_degrid_host(h, B.ptr(grid), B.ptr(vis)))
    ref = O.grid(uv, [1.0], None, geo, 0, grid_in=grid, nthreads=8).reshape(-1)
    assert O.max_relative_error(vis, ref) < tol
    v = g.normal(size=(100000, 2)).view(np.complex128).reshape(-1)
    out = np.zeros((256, 129), dtype=np.complex128)
    B.check(B.lib().lt_nufft_grid_host(h, B.ptr(v), B.ptr(out)))
    refg = O.grid(uv, [1.0], v.reshape(1, -1), geo, 1, nthreads=8)
    assert O.max_relative_error(out, refg) < tol
    # determinism: the gridder reduces partial windows in a fixed order
    out2 = np.zeros((256, 129), dtype=np.complex128)
    B.check(B.lib().lt_nufft_grid_host(h, B.ptr(v), B.ptr(out2)))
    if prec == 0:
        assert np.max(np.abs(out2 - out)) <= 1e-13 * np.max(np.abs(out))
    B.lib().lt_nufft_destroy(h)


@pytest.mark.parametrize('prec,tol', [(0, 1e-12), (1, 1e-5)])
def test_degrid_on_grid_points(hip, oracle, prec, tol):
    """Visibilities exactly on grid lines (fractional offset 0 on an axis: all
    2W+1 taps inside the support) mixed with general ones, in the same tiles
    and waves: the degridder's f == 0 path and its 2W-tap path both match the
    oracle.  On-line coordinates are integer multiples k du whose quotient by
    du (what the plan and the oracle compute) is exactly k."""
    B, O = hip, oracle
    g = rng(67)
    n = 60000
    geo = O.NUFFTGeometry(O.Grid2D(128, 128, [-1.0, 1.0, -1.0, 1.0]), pad=2, wsup=3, beta=14.04)
    uv = np.zeros((n, 3))
    uv[:, 0] = g.uniform(-100.0, 100.0, size=n) * geo.du
    uv[:, 1] = g.uniform(-100.0, 100.0, size=n) * geo.dv
    on_u = g.random(n) < 0.3
    on_v = g.random(n) < 0.3
    ku = np.round(uv[:, 0] / geo.du)
    kv = np.round(uv[:, 1] / geo.dv)
    uv[on_u, 0] = ku[on_u] * geo.du
    uv[on_v, 1] = kv[on_v] * geo.dv
    on_u &= uv[:, 0] / geo.du == ku
    on_v &= uv[:, 1] / geo.dv == kv
    assert np.sum(on_u & on_v) > 1000
    h, keep = _plan(B, uv, [1.0], 128, 128, [(-1.0, 1.0), (-1.0, 1.0)], 3, 14.04, prec)
    grid = np.ascontiguousarray(g.normal(size=(256, 129, 2)).view(np.complex128)[..., 0])
    vis = np.zeros(n, dtype=np.complex128)
    B.check(B.lib().lt_nufft_degrid_host(h, B.ptr(grid), B.ptr(vis)))
    ref = O.grid(uv, [1.0], None, geo, 0, grid_in=grid, nthreads=8).reshape(-1)
    assert O.max_relative_error(vis, ref) < tol
    for sel in (on_u | on_v, ~(on_u | on_v)):
        assert np.max(np.abs(vis[sel] - ref[sel])) < tol * np.max(np.abs(ref))
    B.lib().lt_nufft_destroy(h)


@pytest.mark.parametrize('prec,tol', [(0, 1e-12), (1, 1e-5)])
def test_degrid_forms_agree(hip, oracle, prec, tol):
    """Both tap evaluations of the degridder (lt_nufft_set_degrid_form): the
    window taps in the offset from the cell centre (default) and the
    per-offset taps around the nearest grid point picked by the offset's
    sign, on random visibilities plus a block of visibilities exactly half a
    cell off the grid on one or both axes (the rounding boundary of the
    nearest grid point, offset +-1/2): each matches the oracle and the other."""
    B, O = hip, oracle
    g = rng(69)
    uv = annulus_uv(100000, seed=69)
    du = 1.0 / (4.8481368111e-6 * 5.12)  # the padded c1 image: 256 cells of 0.02"
    half = np.zeros((4096, 3))
    half[:, :2] = (np.floor(g.uniform(-45.0, 45.0, size=(4096, 2))) + 0.5) * du
    half[::2, 1] = g.uniform(-45.0, 45.0, size=2048) * du
    uv = np.concatenate([uv, half])
    n = uv.shape[0]
    h, keep = _plan(B, uv, [1.0], 128, 128, [(-1.28, 1.28), (-1.28, 1.28)], 3, 14.04, prec)
    geo = O.NUFFTGeometry(O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28]), pad=2, wsup=3, beta=14.04)
    grid = np.ascontiguousarray(g.normal(size=(256, 129, 2)).view(np.complex128)[..., 0])
    ref = O.grid(uv, [1.0], None, geo, 0, grid_in=grid, nthreads=8).reshape(-1)
    out = []
    for form in (1, 0):
        B.check(B.lib().lt_nufft_set_degrid_form(h, form))
        vis = np.zeros(n, dtype=np.complex128)
        B.check(B.lib().lt_nufft_degrid_host(h, B.ptr(grid), B.ptr(vis)))
        err = O.max_relative_error(vis, ref)
        print(f'precision {prec} degrid form {form}: {err:.2e} against the oracle')
        assert err < tol
        assert np.max(np.abs(vis[-4096:] - ref[-4096:])) < tol * np.max(np.abs(ref))
        out.append(vis)
    assert O.max_relative_error(out[0], out[1]) < tol
    assert B.lib().lt_nufft_set_degrid_form(h, 2) == B.LT_ERR_ARG
    B.lib().lt_nufft_destroy(h)


def test_grid_cpu_drop_in_symbol(hip, oracle):
    from lentils_amd import FourierSpace, ImageSpace, VisibilitySpace
    B, O = hip, oracle
    uv = annulus_uv(20000, seed=64)
    vs = VisibilitySpace(channels=[1.0], uvcoords=uv)
    fs = FourierSpace(ImageSpace(shape=(256, 256), bounds=[(-2.56, 2.56), (-2.56, 2.56)]))
    geo = O.NUFFTGeometry(O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28]), pad=2, wsup=3, beta=14.04)
    g = rng(65)
    grid = np.ascontiguousarray(g.normal(size=(256, 129, 2)).view(np.complex128)[..., 0])
    vis = g.normal(size=(20000, 2)).view(np.complex128).reshape(-1).copy()
    vis0 = vis.copy()
    B.lib().grid_cpu(vs, B.ptr(vis), fs, B.ptr(grid), 3, 14.04, 0)  # vis += degrid(grid)
    ref = vis0 + O.grid(uv, [1.0], None, geo, 0, grid_in=grid).reshape(-1)
    assert O.max_relative_error(vis, ref) < 1e-12
    grid0 = grid.copy()
    B.lib().grid_cpu(vs, B.ptr(vis0), fs, B.ptr(grid), 3, 14.04, 1)  # grid += grid(vis)
    refg = grid0 + O.grid(uv, [1.0], vis0.reshape(1, -1), geo, 1)
    assert O.max_relative_error(grid, refg) < 1e-12


def test_nufft_adjointness_gpu(hip, oracle):
    B, O = hip, oracle
    uv = annulus_uv(50000, seed=66)
    h, keep = _plan(B, uv, [1.0], 128, 96, [(-1.28, 1.28), (-0.96, 0.96)], 3, 14.04, 0)
    g = rng(67)
    x = g.normal(size=128 * 96)
    y = g.normal(size=(50000, 2)).view(np.complex128).reshape(-1)
    Nx = np.zeros(50000, dtype=np.complex128)
    NTy = np.zeros(128 * 96)
    B.check(B.lib().lt_nufft_forward_host(h, B.ptr(x), B.ptr(Nx)))
    B.check(B.lib().lt_nufft_adjoint_host(h, B.ptr(y), B.ptr(NTy)))
    lhs = np.sum(Nx.real * y.real + Nx.imag * y.imag)
    rhs = np.dot(x, NTy)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))
    B.lib().lt_nufft_destroy(h)


def test_fft_operator(hip, oracle):
    """FFTOperator (fourier_operators.py:12-24) on rocFFT: forward equals
    numpy's rfft2 (unnormalised), transpose numpy's irfft2 with
    norm='forward', including spectra that are not Hermitian in columns 0
    and ny/2 (numpy reads only their Hermitian part); odd and even sizes."""
    O = oracle
    from lentils_amd import FFTOperator, ImageSpace
    for nx, ny in ((64, 48), (33, 50), (40, 31)):
        im = ImageSpace(shape=(nx, ny), bounds=[(-1, 1), (-1, 1)])
        op = FFTOperator(im)
        x = rng(46).normal(size=(nx, ny))
        X = op * x
        ref = np.fft.rfft2(x, norm='backward')
        assert O.max_relative_error(X.reshape(nx, ny // 2 + 1), ref) < 1e-13
        g = rng(47).normal(size=(nx, ny // 2 + 1, 2)).view(np.complex128)[..., 0]
        y = op.T * g
        refT = np.fft.irfft2(g, s=(nx, ny), norm='forward')
        assert O.max_relative_error(y.reshape(nx, ny), refT) < 1e-13


def test_grid_cpu_wraps_beyond_the_band(hip, oracle):
    """grid_cpu on visibilities whose taps cross the grid's edge: the
    reference wraps them periodically (nufft.c:83-86), so does the drop-in
    (a direct per-visibility pass when the tiled gridder's band check
    refuses the uv set); both directions against the oracle."""
    from lentils_amd import FourierSpace, ImageSpace, VisibilitySpace
    B, O = hip, oracle
    uv = annulus_uv(6000, r1=4.4e6, r2=5.6e6, seed=68)
    vs = VisibilitySpace(channels=[1.0], uvcoords=uv)
    fs = FourierSpace(ImageSpace(shape=(64, 64), bounds=[(-0.64, 0.64), (-0.64, 0.64)]))
    geo = O.NUFFTGeometry(O.Grid2D(32, 32, [-0.32, 0.32, -0.32, 0.32]), pad=2, wsup=3, beta=14.04)
    assert np.max(np.abs(uv[:, :2])) / geo.du + 3 + 1 > 32  # the Nyquist guard would refuse these
    g = rng(69)
    grid = np.ascontiguousarray(g.normal(size=(64, 33, 2)).view(np.complex128)[..., 0])
    vis = g.normal(size=(6000, 2)).view(np.complex128).reshape(-1).copy()
    vis0 = vis.copy()
    B.lib().grid_cpu(vs, B.ptr(vis), fs, B.ptr(grid), 3, 14.04, 0)
    ref = vis0 + O.grid(uv, [1.0], None, geo, 0, grid_in=grid).reshape(-1)
    assert O.max_relative_error(vis, ref) < 1e-12
    grid0 = grid.copy()
    B.lib().grid_cpu(vs, B.ptr(vis0), fs, B.ptr(grid), 3, 14.04, 1)
    refg = grid0 + O.grid(uv, [1.0], vis0.reshape(1, -1), geo, 1)
    assert O.max_relative_error(grid, refg) < 1e-12
