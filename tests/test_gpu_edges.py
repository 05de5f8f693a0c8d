"""Edge cases through the C ABI, against the oracle: no points, no
visibilities, no lens components (the identity map), a fully masked image,
one-pixel and one-row source grids, visibilities whose kernel taps end
exactly at the Nyquist band's edge."""
import ctypes as C

import numpy as np
import pytest

from helpers import c1_lens, rng
from test_gpu_lens import _glm, _space

pytestmark = pytest.mark.gpu


def test_no_points(hip, oracle):
    B, O = hip, oracle
    glm, keep = _glm(B, c1_lens(O), O)
    pin = np.zeros((1, 2))
    pout = np.full((1, 2), 7.0)
    B.lib().deflect_points(glm, B.ptr(pin), 0, B.ptr(pout), 0, None)
    B.check(B.lib().lt_deflect_points(glm, B.ptr(pin), 0, B.ptr(pout)))
    assert np.all(pout == 7.0)


def test_identity_lens_and_full_mask(hip, oracle):
    """No lens components: every image pixel maps to itself, each row the
    clipped areas of its 8 fan triangles after the reference's reduction (the
    source grid shifted by a non-dyadic fraction of a pixel so that no
    stencil point lies on a grid line); and an image whose every pixel is
    masked gives no entries."""
    B, O = hip, oracle
    im = O.Grid2D(24, 20, [-0.48, 0.48, -0.4, 0.4])
    src = O.Grid2D(16, 16, [-0.3013, 0.3387, -0.2971, 0.3429])
    arr = np.zeros(0, dtype=O.MASS_MODEL_DTYPE)
    glm = B.global_lens_model_ctype(0, 1, arr.ctypes.data_as(C.c_void_p))
    n = im.nx * im.ny
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(50 * n, dtype=np.int32)
    va = np.zeros(50 * n)
    B.lib().manifold_lens_matrix_csr(_space(B, im), _space(B, src), glm, B.ptr(rp), B.ptr(ci), B.ptr(va))
    orp, oci, ova, _ = O.lens_matrix_csr(im, src, [])
    assert np.array_equal(rp, orp) and np.array_equal(ci[:rp[-1]], oci)
    assert np.max(np.abs(va[:rp[-1]] - ova)) <= 1e-12
    # every pixel has the same offset to the source grid: the interior rows
    # (all 8 triangles inside the grid) have one and the same sum
    sums = np.add.reduceat(va[:rp[-1]], rp[:-1][np.diff(rp) > 0])
    top = sums.max()
    assert np.sum(np.abs(sums - top) < 1e-12) >= 100
    im.mask[:] = False
    rp2 = np.full(n + 1, -1, dtype=np.int32)
    B.lib().manifold_lens_matrix_csr(_space(B, im), _space(B, src), glm, B.ptr(rp2), B.ptr(ci), B.ptr(va))
    orp2, oci2, ova2, _ = O.lens_matrix_csr(im, src, [])
    assert np.array_equal(rp2, orp2)
    assert np.array_equal(ci[:rp2[-1]], oci2) and np.array_equal(va[:rp2[-1]], ova2)


@pytest.mark.parametrize('shape', [(1, 1), (1, 7), (6, 1)])
def test_curvature_of_thin_grids(hip, oracle, shape):
    B, O = hip, oracle
    g = O.Grid2D(shape[0], shape[1], [-0.1, 0.1 * shape[0], -0.1, 0.1 * shape[1]])
    n = g.nx * g.ny
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(5 * n, dtype=np.int32)
    va = np.zeros(5 * n)
    B.lib().curv_matrix_findiff_csr(_space(B, g), B.ptr(rp), B.ptr(ci), B.ptr(va))
    orp, oci, ova = O.curv_matrix(g)
    assert np.array_equal(rp, orp) and np.array_equal(ci[:rp[-1]], oci) and np.array_equal(va[:rp[-1]], ova)


def test_no_visibilities(hip, oracle):
    """grid_cpu and dft_matrix_csr with an empty uv set leave the outputs as
    they are; the NUFFT plan maps to and from an empty visibility set."""
    from lentils_amd import FourierSpace, ImageSpace, NUFFTOperator, VisibilitySpace
    B, O = hip, oracle
    uv = np.zeros((0, 3))
    vs = VisibilitySpace(channels=[1.0], uvcoords=uv)
    ims = ImageSpace(shape=(32, 32), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
    fs = FourierSpace(ImageSpace(shape=(64, 64), bounds=[(-1.28, 1.28), (-1.28, 1.28)]))
    grid = np.ascontiguousarray(rng(3).normal(size=(64, 33, 2)).view(np.complex128)[..., 0])
    g0 = grid.copy()
    vis = np.zeros(1, dtype=np.complex128)
    B.lib().grid_cpu(vs, B.ptr(vis), fs, B.ptr(grid), 3, 14.04, 1)
    B.lib().grid_cpu(vs, B.ptr(vis), fs, B.ptr(grid), 3, 14.04, 0)
    assert np.array_equal(grid, g0)
    rp = np.full(1, 5, dtype=np.int32)
    ci = np.zeros(1, dtype=np.int32)
    va = np.zeros(1)
    g = O.Grid2D(32, 32, [-0.64, 0.64, -0.64, 0.64])
    B.lib().dft_matrix_csr(vs, _space(B, g), B.ptr(rp), B.ptr(ci), B.ptr(va))
    assert rp[0] == 0
    # the operator class refuses it as the reference's does (its Nyquist
    # guard takes np.max of the empty uv set, fourier_operators.py:77-78)
    with pytest.raises(ValueError):
        NUFFTOperator(vs, ims, kernel_support=3)
    # the plan itself maps to and from the empty set
    from lentils_amd.solvers import NUFFTPlan
    plan = NUFFTPlan(vs, ims, 2, 3, 14.04)
    assert plan.forward(rng(4).normal(size=(32, 32))).size == 0
    back = plan.adjoint(np.zeros(0, dtype=np.complex128))
    assert back.size == 32 * 32 and np.all(back == 0)


def test_taps_at_the_band_edge(hip, oracle):
    """Visibilities whose outermost kernel tap lands on the last cell the
    Nyquist guard admits (|u| / du + W + 1 = nu / 2 at most,
    fourier_operators.py:77-80): forward and adjoint against the oracle,
    FP64."""
    from lentils_amd import ImageSpace, VisibilitySpace
    from lentils_amd.solvers import NUFFTPlan
    B, O = hip, oracle
    n = 64
    ims = ImageSpace(shape=(n, n), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
    geo = O.NUFFTGeometry(O.Grid2D(n, n, [-0.64, 0.64, -0.64, 0.64]), pad=2, wsup=3, beta=14.04)
    edge = geo.nu // 2 - 3 - 1  # the largest admitted |ngp|
    g = rng(8)
    m = 400
    cells = g.integers(-edge, edge + 1, size=(m, 2))
    cells[: m // 4, 0] = edge * g.choice([-1, 1], m // 4)
    cells[m // 4: m // 2, 1] = edge * g.choice([-1, 1], m // 4)
    frac = g.uniform(-0.49, 0.49, size=(m, 2))
    at = np.abs(cells) == edge
    frac[at] = -np.sign(cells[at]) * np.abs(frac[at])  # |u| / du <= edge: the reference's guard admits it too
    uv = np.zeros((m, 3))
    uv[:, 0] = (cells[:, 0] + frac[:, 0]) * geo.du
    uv[:, 1] = -(cells[:, 1] + frac[:, 1]) * geo.dv  # v = -V
    vs = VisibilitySpace(channels=[1.0], uvcoords=uv)
    plan = NUFFTPlan(vs, ims, 2, 3, 14.04)
    img = g.normal(size=(n, n))
    ref = O.nufft_forward(img, uv, [1.0], geo).reshape(-1)
    assert O.max_relative_error(plan.forward(img).reshape(-1), ref) < 1e-12
    vis = g.normal(size=(m, 2)).view(np.complex128).reshape(-1)
    refa = O.nufft_adjoint(vis.reshape(1, -1), uv, [1.0], geo)
    assert O.max_relative_error(plan.adjoint(vis).reshape(n, n), refa) < 1e-12
