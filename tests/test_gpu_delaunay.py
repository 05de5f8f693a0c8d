"""GPU Delaunay source planes (triangles.c, lens_operators.py:45-84) and the
reference's outer optimisation loop (optimization_tests.py:90-165):

* the Delaunay lens matrix and the triangle gradient rows are bit-identical
  to the oracle (itself bit-equal to the reference build) on the same
  inputs; triangle_geometry_from_inds likewise;
* DelaunayLensOperator end to end: GPU deflections within 1e-12 of the
  oracle, its matrix equal to the oracle's assembly from its own inputs;
* PriorCovarianceOperator on a DelaunaySpace ('gradient', boundary term);
* three iterations of the log-normal Gauss-Newton loop of
  test_lognormal_radio on the reference's uvfits file: every CG solve meets
  its tolerance and the data misfit falls."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from helpers import c1_lens, rng
from test_oracle_delaunay import delaunay_inputs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TINY = os.path.join(HERE, 'golden', 'j0751_lwmp_nopos_nosub_0.000000_tiny.uvfits')


def _dspace(B, pts, tris):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    ds = B.delaunay_space_ctype(pts.shape[0], tris.shape[0], 1, 1, None, pts.ctypes.data_as(B.c_double_p),
                                tris.ctypes.data_as(B.c_int_p))
    return ds, (pts, tris)


@pytest.mark.parametrize('ncast,masked', [(1, False), (2, True), (3, True)])
def test_delaunay_matrix_bit_exact(hip, oracle, ncast, masked):
    B, O = hip, oracle
    n = 96
    mask = None
    if masked:
        g = np.indices((n, n))
        r = np.hypot(g[0] - n / 2, g[1] - n / 2)
        mask = (r > 4) & (r < 0.45 * n)
    im = O.Grid2D(n, n, [-1.28, 1.28, -1.28, 1.28], mask=mask)
    pts, tris, up, ut = delaunay_inputs(O, im, c1_lens(O), ncast, mask)
    rp, ci, va = O.delaunay_lens_matrix_csr(im, pts, tris, ncast, up, ut)
    ds, keep = _dspace(B, pts, tris)
    ims = B.image_space_ctype(n, n, 1, 1, im.xmin, im.xmax, im.ymin, im.ymax, im.dx, im.dy, None,
                              im.mask.ctypes.data_as(C.POINTER(C.c_bool)))
    grp = np.zeros(n * n + 1, dtype=np.int32)
    gci = np.zeros(ci.size + 1, dtype=np.int32)
    gva = np.zeros(ci.size + 1)
    nnz = C.c_long(0)
    upc, utc = np.ascontiguousarray(up), np.ascontiguousarray(ut)
    B.check(B.lib().lt_delaunay_lens_matrix_csr(ims, ds, ncast, B.ptr(upc), B.ptr(utc), B.ptr(grp), B.ptr(gci),
                                                B.ptr(gva), gci.size, C.byref(nnz)))
    assert nnz.value == rp[-1]
    assert np.array_equal(grp, rp) and np.array_equal(gci[:nnz.value], ci) and np.array_equal(gva[:nnz.value], va)
    # the reference symbol itself
    drp = np.zeros_like(grp)
    dci = np.zeros_like(gci)
    dva = np.zeros_like(gva)
    B.lib().delaunay_lens_matrix_csr(ims, ds, ncast, B.ptr(upc), B.ptr(utc), B.ptr(drp), B.ptr(dci), B.ptr(dva))
    assert np.array_equal(drp, rp) and np.array_equal(dva[:rp[-1]], va)
    # gradient rows
    orp, oci, ova = O.triangle_gradient_csr(pts, tris)
    nt = tris.shape[0]
    trp = np.zeros(2 * nt + 1, dtype=np.int32)
    tci = np.zeros(6 * nt, dtype=np.int32)
    tva = np.zeros(6 * nt)
    B.lib().triangle_gradient_csr(ds, B.ptr(trp), B.ptr(tci), B.ptr(tva))
    assert np.array_equal(trp, orp) and np.array_equal(tci, oci) and np.array_equal(tva, ova)


def test_triangle_geometry_from_inds(hip, oracle):
    B, O = hip, oracle
    from scipy.spatial import Delaunay
    g = rng(72)
    P = np.ascontiguousarray(g.uniform(-1, 1, size=(500, 2)))
    tri = Delaunay(P)
    s = g.uniform(-0.6, 0.6, size=(5000, 2))
    t = tri.find_simplex(s)
    s = np.ascontiguousarray(s[t >= 0])
    idx = np.ascontiguousarray(tri.simplices[t[t >= 0]], dtype=np.int32)
    w = np.zeros((s.shape[0], 3))
    B.lib().triangle_geometry_from_inds(B.ptr(s), B.ptr(idx), B.ptr(w), s.shape[0], B.ptr(np.ascontiguousarray(tri.points)))
    wo = np.zeros_like(w)
    O.lib().oracle_triangle_geometry_from_inds(O._ptr(s), O._ptr(idx), O._ptr(wo), s.shape[0],
                                               O._ptr(np.ascontiguousarray(tri.points)))
    assert np.array_equal(w, wo)


def test_delaunay_lens_operator_and_prior(hip, oracle):
    O = oracle
    from lentils_amd import DelaunayLensOperator, ExternalPotential, GlobalLensModel, ImageSpace, \
        PowerLawEllipsoid, PriorCovarianceOperator
    n = 80
    g = np.indices((n, n))
    r = np.hypot(g[0] - n / 2, g[1] - n / 2)
    mask = (r > 3) & (r < 0.45 * n)
    im = ImageSpace(shape=(n, n), bounds=[(-1.28, 1.28), (-1.28, 1.28)], mask=mask)
    lm = GlobalLensModel([PowerLawEllipsoid(b=1.0, f=0.8, th=30.0, qh=0.5, x=0.02, y=-0.01, rc=0.0, z=0.5),
                          ExternalPotential(ss=0.03, sa=45.0, x=0.02, y=-0.01, z=0.5)])
    op = DelaunayLensOperator(im, lm, z_src=2.0, ncasted=2)
    src = op.space_right
    cast = np.zeros((n, n), dtype=bool)
    cast[::2, ::2] = True
    cp_ref = O.deflect_points(c1_lens(O), np.ascontiguousarray(im.points[cast & mask]))
    a_g, a_r = im.points[cast & mask] - op.casted_points, im.points[cast & mask] - cp_ref
    assert np.max(np.linalg.norm(a_g - a_r, axis=1) / np.linalg.norm(a_r, axis=1)) < 1e-12
    og = O.Grid2D(n, n, im.bounds, mask=mask)
    ut = src.delaunay.find_simplex(op.uncasted_points).astype(np.int32)
    rp, ci, va = O.delaunay_lens_matrix_csr(og, src.points, src.triangles, 2, op.uncasted_points, ut)
    M = op.matrix.tocsr()
    assert np.array_equal(M.indptr, rp) and np.array_equal(M.indices, ci) and np.array_equal(M.data, va)
    x = rng(73).normal(size=src.shape)
    assert O.max_relative_error((op * x).reshape(-1), M @ x.reshape(-1)) < 1e-13
    pr = PriorCovarianceOperator(src, type='gradient', strength=3.0, lambda_boundary=10.0)
    grp, gci, gva = O.triangle_gradient_csr(src.points, src.triangles)
    G = sp.csr_matrix((gva, gci, grp), shape=(2 * src.num_triangles, src.size))
    bnd = np.zeros(src.size)
    bnd[src.edge] = 3.0 * 10.0
    ref = 3.0 * (G.T @ (G @ x.reshape(-1))) + bnd * x.reshape(-1)
    assert O.max_relative_error((pr * x).reshape(-1), ref) < 1e-12


def test_lognormal_radio_outer_loop(hip, oracle):
    """optimization_tests.py:90-165 on the reference's tiny j0751 file,
    shortened: log-normal source s = exp(z) on a Delaunay source plane,
    Gauss-Newton steps solving (J^T L^T BCB L J + R) dz = -(J^T L^T (BCB L s - d)
    + R z) by CG with the reference's splu(R + wtsum J^T L^T L J)
    preconditioner; each solve to the reference's tolerance, the misfit
    falling."""
    from lentils_amd import DelaunayLensOperator, DiagonalOperator, ExternalPotential, GlobalLensModel, \
        ImageSpace, PowerLawEllipsoid, PriorCovarianceOperator, RadioDataset
    # 128^2 pixels of 0.0015" (the file's 6e7-wavelength baselines need
    # <= 0.0016") on the Einstein ring east of the lens centre
    n = 128
    im = ImageSpace(shape=(n, n), bounds=[(-0.096, 0.096), (0.079, 0.271)])
    g = np.indices((n, n))
    r = np.hypot(g[0] - n / 2, g[1] - n / 2)
    im.mask = r < 0.45 * n
    im._refresh()
    uv = RadioDataset(TINY, image_space=im)
    bcb = uv.blurred_covariance_operator
    dirty = uv.dirty_image
    lm = GlobalLensModel([PowerLawEllipsoid(b=0.462437, th=19.162881, f=0.899243, x=-0.441519, y=0.175067, rc=0.0,
                                            qh=0.449087, z=0.35),
                          ExternalPotential(ss=0.092515, sa=73.698888, z=0.35)])
    lens = DelaunayLensOperator(im, lm, z_src=2.059, ncasted=2)
    src = lens.space_right
    wtsum = float(np.sum(np.abs(uv.covariance_operator.T * uv.space.new_vector(1.0))))
    z = src.new_vector(-10.0)
    s = np.exp(z)
    misfit = []
    L = lens.matrix.tocsr()
    for it in range(3):
        lams = 1e3 * 2 ** (3 - it)
        reg = PriorCovarianceOperator(src, type='gradient', strength=lams)
        jz = DiagonalOperator(src, s)
        ddata = bcb * (lens * s) - dirty
        misfit.append(float(np.sum((lens * s).reshape(-1) * (bcb * (lens * s)).reshape(-1))
                            - 2 * np.sum((lens * s).reshape(-1) * dirty.reshape(-1))))
        lhs = jz.T * lens.T * bcb * lens * jz + reg
        rhs = -(jz.T * lens.T * ddata + reg * z)
        J = sp.diags(s.reshape(-1))
        lu = spl.splu((reg._mat + wtsum * J @ L.T @ L @ J).tocsc())
        A = spl.LinearOperator((src.size, src.size), matvec=lambda v: lhs.apply(v.reshape(src.shape)).reshape(-1),
                               dtype=np.float64)
        M = spl.LinearOperator((src.size, src.size), matvec=lu.solve, dtype=np.float64)
        sol, info = spl.cg(A, rhs.reshape(-1), rtol=1e-10, atol=1e-14, M=M, maxiter=2000)
        res = A @ sol - rhs.reshape(-1)
        assert info == 0 and np.max(np.abs(res)) / np.max(np.abs(rhs)) < 1e-6
        z = z + sol.reshape(src.shape)
        s = np.exp(z)
    final = float(np.sum((lens * s).reshape(-1) * (bcb * (lens * s)).reshape(-1))
                  - 2 * np.sum((lens * s).reshape(-1) * dirty.reshape(-1)))
    print('misfit', misfit + [final])
    assert final < misfit[0]
