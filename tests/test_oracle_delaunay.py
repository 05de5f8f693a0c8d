"""Pins the oracle's Delaunay source-plane functions (triangles.c) on the
CPU: bit for bit against the reference's own triangles.c (oracle/_ref),
and by what the mathematics fixes: barycentric weights sum to 1 and
reproduce the point, and the gradient rows applied to a linear function
give its gradient times sqrt|area| (times the sign of the triangle's
orientation)."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

from helpers import c1_lens, rng


def delaunay_inputs(O, im, lenses, ncast, mask=None):
    """What DelaunayLensOperator (lens_operators.py:45-84) computes before
    calling the C: casted / uncasted pixels, their deflections, the
    triangulation of the casted ones and the uncasted points' triangles."""
    pts = np.stack(np.meshgrid(im.xmin + im.dx * np.arange(im.nx), im.ymin + im.dy * np.arange(im.ny),
                               indexing='ij'), -1)
    m = np.ones((im.nx, im.ny), dtype=bool) if mask is None else mask
    cast = np.zeros_like(m)
    cast[::ncast, ::ncast] = True
    unc = ~cast
    cast &= m
    unc &= m
    cp = O.deflect_points(lenses, np.ascontiguousarray(pts[cast]))
    up = O.deflect_points(lenses, np.ascontiguousarray(pts[unc]))
    tri = Delaunay(cp)
    ut = tri.find_simplex(up).astype(np.int32)
    return tri.points, tri.simplices.astype(np.int32), up, ut


def test_delaunay_oracle_matches_reference_build(oracle):
    O = oracle
    if O.ref_lib() is None:
        pytest.skip('reference sources absent: oracle/_ref not built')
    lenses = c1_lens(O)
    for ncast, n, masked in ((1, 40, False), (2, 48, True), (3, 45, True)):
        mask = None
        if masked:
            g = np.indices((n, n))
            r = np.hypot(g[0] - n / 2, g[1] - n / 2)
            mask = (r > 3) & (r < 0.45 * n)
        im = O.Grid2D(n, n, [-1.28, 1.28, -1.28, 1.28], mask=mask)
        pts, tris, up, ut = delaunay_inputs(O, im, lenses, ncast, mask)
        assert np.any(ut < 0) or ncast == 1
        a = O.delaunay_lens_matrix_csr(im, pts, tris, ncast, up, ut)
        b = O.delaunay_lens_matrix_csr(im, pts, tris, ncast, up, ut, use_ref=True)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        ga = O.triangle_gradient_csr(pts, tris)
        gb = O.triangle_gradient_csr(pts, tris, use_ref=True)
        for x, y in zip(ga, gb):
            assert np.array_equal(x, y)


def test_barycentric_weights_reproduce_points(oracle):
    O = oracle
    g = rng(70)
    P = g.uniform(-1, 1, size=(300, 2))
    tri = Delaunay(P)
    s = g.uniform(-0.5, 0.5, size=(2000, 2))
    t = tri.find_simplex(s)
    s, t = s[t >= 0], t[t >= 0].astype(np.int32)
    idx = np.ascontiguousarray(tri.simplices[t], dtype=np.int32)
    w = np.zeros((s.shape[0], 3))
    O.lib().oracle_triangle_geometry_from_inds(O._ptr(np.ascontiguousarray(s)), O._ptr(idx), O._ptr(w), s.shape[0],
                                               O._ptr(np.ascontiguousarray(tri.points)))
    assert np.allclose(w.sum(axis=1), 1.0, atol=1e-13)
    assert np.all(w > -1e-12)
    rec = np.einsum('nk,nkd->nd', w, tri.points[idx])
    assert np.allclose(rec, s, atol=1e-13)


def test_triangle_gradient_rows_on_linear_functions(oracle):
    O = oracle
    g = rng(71)
    P = g.uniform(-1, 1, size=(200, 2))
    tri = Delaunay(P)
    rp, ci, va = O.triangle_gradient_csr(tri.points, tri.simplices)
    import scipy.sparse as sp
    G = sp.csr_matrix((va, ci, rp), shape=(2 * tri.simplices.shape[0], P.shape[0]))
    a, b = 0.7, -1.9
    f = a * tri.points[:, 0] + b * tri.points[:, 1]
    v = tri.points[tri.simplices]
    area = 0.5 * ((v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) -
                  (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1]))
    out = (G @ f).reshape(-1, 2)
    s = np.sqrt(np.abs(area)) * np.sign(area)
    assert np.allclose(out[:, 0], a * s, rtol=1e-10, atol=1e-13)
    assert np.allclose(out[:, 1], b * s, rtol=1e-10, atol=1e-13)
    assert np.allclose(G @ np.ones(P.shape[0]), 0.0, atol=1e-12)
