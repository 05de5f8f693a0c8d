"""Pins the oracle's NUFFT chain (CPU, no GPU).

nufft.c's grid_cpu cannot be built here (it needs GSL), so the oracle's
gridding is pinned by what the mathematics fixes:
* against the brute-force DFT (dft_matrix_csr restated), the reference's own
  test nufft_tests.py:43-67 at its sizes and its 1e-6 bar;
* adjointness <N x, y> = <x, N^T y> in the real inner product;
* I0 against scipy.special.i0; the kernel's support and normalisation.
GSL's exact I0 bits remain unpinned (tolerances do not depend on them).
"""
import numpy as np
import scipy.sparse as sp
import scipy.special as sps

from helpers import reference_test_uv, rng


def _dft(O, uv, ch, im):
    rp, ci, va = O.dft_matrix(uv, ch, im)
    return sp.csr_matrix((va, ci, rp), shape=(2 * uv.shape[0], im.nx * im.ny))


def test_nufft_matches_dft_reference_test(oracle):
    O = oracle
    uv = reference_test_uv()
    ch = [0.5]
    im = O.Grid2D(236, 257, [-0.5, 0.7, -0.7, 0.5])
    geo = O.NUFFTGeometry(im, pad=2, wsup=4)
    D = _dft(O, uv, ch, im)
    v = rng(9).normal(size=(uv.shape[0], 2)).view(np.complex128).reshape(-1)
    dftvec = (D.T @ v.view(np.float64)).reshape(236, 257)
    nuvec = O.nufft_adjoint(v.reshape(1, -1), uv, ch, geo)
    assert O.max_relative_error(nuvec, dftvec) < 1e-6
    nuvis = O.nufft_forward(dftvec, uv, ch, geo).reshape(-1)
    dftvis = (D @ dftvec.reshape(-1)).view(np.complex128)
    assert O.max_relative_error(nuvis, dftvis) < 1e-6


def test_nufft_adjointness(oracle):
    O = oracle
    uv = reference_test_uv(seed=8)
    im = O.Grid2D(64, 48, [-0.3, 0.34, -0.2, 0.28])
    for wsup, beta in [(4, None), (3, 14.04)]:
        geo = O.NUFFTGeometry(im, pad=2, wsup=wsup, beta=beta)
        g = rng(10)
        x = g.normal(size=(64, 48))
        y = g.normal(size=(uv.shape[0], 2)).view(np.complex128).reshape(-1)
        Nx = O.nufft_forward(x, uv, [1.0], geo).reshape(-1)
        NTy = O.nufft_adjoint(y.reshape(1, -1), uv, [1.0], geo)
        lhs = np.sum(Nx.real * y.real + Nx.imag * y.imag)
        rhs = np.sum(x * NTy)
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


def test_bessel_i0_and_kernel(oracle):
    O = oracle
    xs = np.linspace(0, 40, 2001)
    ours = np.array([O.bessel_i0(x) for x in xs])
    assert np.max(np.abs(ours / sps.i0(xs) - 1)) < 5e-15
    beta, w = 14.04, 3.0
    assert O.lib().oracle_kb_g(0.0, 0.0, beta, w) == 1.0 / (4 * w * w) * O.bessel_i0(beta) ** 2
    assert O.lib().oracle_kb_g(3.0001, 0.0, beta, w) == 0.0
    assert O.lib().oracle_kb_g(3.0, 0.0, beta, w) == 1.0 / (4 * w * w) * O.bessel_i0(beta)


def test_curvature_matrix_is_laplacian(oracle):
    O = oracle
    im = O.Grid2D(5, 4, [-1, 1, -0.8, 0.8])
    rp, ci, va = O.curv_matrix(im)
    M = sp.csr_matrix((va, ci, rp), shape=(20, 20)).toarray()
    h = 1.0 / (im.dx * im.dy)
    assert np.allclose(M, M.T)
    assert np.allclose(np.diag(M), -4 * h)
    x = rng(3).normal(size=(5, 4))
    lap = -4 * x
    lap[1:] += x[:-1]
    lap[:-1] += x[1:]
    lap[:, 1:] += x[:, :-1]
    lap[:, :-1] += x[:, 1:]
    assert np.allclose(M @ x.reshape(-1), h * lap.reshape(-1))


def test_gradient_matrix_is_forward_difference(oracle):
    """grad_matrix_findiff_csr (nufft.c:163-194): x-differences then
    y-differences, 1/dx and 1/dy, no entries on the last row / column; a
    linear image has the exact constant gradient."""
    O = oracle
    im = O.Grid2D(6, 5, [-1.2, 0.9, -0.7, 0.8])
    rp, ci, va = O.grad_matrix(im)
    G = sp.csr_matrix((va, ci, rp), shape=(60, 30))
    x = rng(4).normal(size=(6, 5))
    gx = np.zeros((6, 5))
    gy = np.zeros((6, 5))
    gx[:-1] = (x[1:] - x[:-1]) / im.dx
    gy[:, :-1] = (x[:, 1:] - x[:, :-1]) / im.dy
    assert np.allclose(G @ x.reshape(-1), np.concatenate([gx.reshape(-1), gy.reshape(-1)]), rtol=0, atol=1e-12)
    a, b = 0.7, -1.3
    lin = a * im.dx * np.arange(6)[:, None] + b * im.dy * np.arange(5)[None, :]
    d = (G @ lin.reshape(-1)).reshape(2, 6, 5)
    assert np.allclose(d[0, :-1], a, atol=1e-12) and np.all(d[0, -1] == 0)
    assert np.allclose(d[1, :, :-1], b, atol=1e-12) and np.all(d[1, :, -1] == 0)
    assert rp[-1] == 2 * (5 * 5 + 6 * 4)


def _bcb_problem(O, n=32, nv=3000, seed=3):
    from helpers import annulus_uv
    im = O.Grid2D(n, n, (-0.64, 0.64, -0.64, 0.64))
    uvw = annulus_uv(nv, r2=1.5e6, seed=seed)
    w = rng(seed + 1).uniform(0.5e6, 1.5e6, nv)
    return im, uvw, w


def test_bcb_oracle_matches_explicit_and_dft(oracle):
    """The dirty-beam convolution (data_base.py:115-129,163-172) equals the
    explicit N^T C N and the DFT's D^T C D to the reference's 1e-6 at its
    default kernel support 4 (nufft_tests.py:87-97); at support 6 both are
    exact to ~1e-11, which a wrong roll, a missing 1/N or an apodization
    term would break."""
    O = oracle
    im, uvw, w = _bcb_problem(O)
    rp, ci, va = O.dft_matrix(uvw, [1.0], im)
    D = sp.csr_matrix((va, ci, rp), shape=(2 * uvw.shape[0], im.nx * im.ny))
    x = rng(5).normal(size=(im.nx, im.ny))
    dft = (D.T @ (np.repeat(w, 2) * (D @ x.reshape(-1)))).reshape(im.nx, im.ny)
    for wsup, bar in ((4, 1e-6), (6, 1e-10)):
        geo = O.NUFFTGeometry(im, pad=2, wsup=wsup)
        beam = O.dirty_beam(uvw, [1.0], im, w, wsup=wsup)
        bcb = O.bcb_apply(x, beam, geo)
        vis = O.nufft_forward(x, uvw, [1.0], geo).reshape(-1) * w
        expl = O.nufft_adjoint(vis.reshape(1, -1), uvw, [1.0], geo)
        assert O.max_relative_error(bcb, expl) < bar
        assert O.max_relative_error(bcb, dft) < bar


def test_dirty_beam_peak_is_weight_sum(oracle):
    """nufft_tests.py:116-118: max(beam) / sum(w) = 1 within 1e-6; the peak
    sits at the beam space's centre pixel (offset 0) and the beam is even."""
    O = oracle
    im, uvw, w = _bcb_problem(O, n=24)
    beam = O.dirty_beam(uvw, [1.0], im, w)
    assert abs(beam.max() / w.sum() - 1.0) < 1e-6
    assert np.unravel_index(np.argmax(beam), beam.shape) == (im.nx, im.ny)
    inner = beam[1:, 1:]
    assert O.max_relative_error(inner, inner[::-1, ::-1]) < 1e-6


def test_grid_wrap_is_periodic_aliasing(oracle):
    """nufft.c:83-93: taps past the grid's edge wrap periodically (and fold
    into the stored half plane).  On a centred image (no phase term) a
    visibility one grid period away, u +- nu du or v +- nv dv, reads and
    writes the same cells with the same kernel weights and checkerboard signs,
    so degridding and gridding agree with the unshifted visibility; a clamp,
    a wrong modulus or a sign taken after the wrap breaks this."""
    O = oracle
    geo = O.NUFFTGeometry(O.Grid2D(32, 32, [-0.32, 0.32, -0.32, 0.32]), pad=2, wsup=3, beta=14.04)
    assert geo.gcx == 0.0 and geo.gcy == 0.0
    g = rng(70)
    n = 400
    uv = np.zeros((n, 3))
    uv[:, 0] = g.uniform(-0.5, 0.5, n) * geo.nu * geo.du
    uv[:, 1] = g.uniform(-0.5, 0.5, n) * geo.nv * geo.dv
    grid = np.ascontiguousarray(g.normal(size=(geo.nu, geo.half_nv, 2)).view(np.complex128)[..., 0])
    vis = g.normal(size=(1, n, 2)).view(np.complex128)[..., 0]
    ref_v = O.grid(uv, [1.0], None, geo, 0, grid_in=grid)
    ref_g = O.grid(uv, [1.0], vis, geo, 1)
    for du, dv in ((1, 0), (-1, 0), (0, 1), (0, -1), (1, -1)):
        sh = uv.copy()
        sh[:, 0] += du * geo.nu * geo.du
        sh[:, 1] += dv * geo.nv * geo.dv  # v = -V: a whole period either way
        got_v = O.grid(sh, [1.0], None, geo, 0, grid_in=grid)
        got_g = O.grid(sh, [1.0], vis, geo, 1)
        assert O.max_relative_error(got_v, ref_v) < 1e-11, (du, dv)
        assert O.max_relative_error(got_g, ref_g) < 1e-11, (du, dv)
