"""Pins the oracle's rasterization (CPU, no GPU).

* bit-for-bit against the reference's own raster.c + deflect.c build, on
  lensed grids with masks, empty rows and the lens-centre pixel;
* geometry against brute-force polygon clipping: without the reference's
  dropped-first-fragment quirk each CSR entry must equal
  sum_t 0.25 |T_t cap P| / |T_t| over the 8 fan triangles of the pixel;
* invariants: an interior pixel of an identity mapping sums to 2.0 before
  the drop; the drop removes exactly one fragment per later column group.
"""
import numpy as np
import pytest

from helpers import c1_lens, rng


def _clip(poly, xmin, xmax, ymin, ymax):
    """Sutherland-Hodgman clip of a convex polygon to an axis box."""
    def clip_edge(pts, inside, inter):
        out = []
        n = len(pts)
        for k in range(n):
            a, b = pts[k], pts[(k + 1) % n]
            ia, ib = inside(a), inside(b)
            if ia:
                out.append(a)
                if not ib:
                    out.append(inter(a, b))
            elif ib:
                out.append(inter(a, b))
        return out

    def ix(x0):
        return lambda a, b: (x0, a[1] + (b[1] - a[1]) * (x0 - a[0]) / (b[0] - a[0]))

    def iy(y0):
        return lambda a, b: (a[0] + (b[0] - a[0]) * (y0 - a[1]) / (b[1] - a[1]), y0)

    pts = list(poly)
    for inside, inter in [(lambda p: p[0] >= xmin, ix(xmin)), (lambda p: p[0] <= xmax, ix(xmax)),
                          (lambda p: p[1] >= ymin, iy(ymin)), (lambda p: p[1] <= ymax, iy(ymax))]:
        if not pts:
            return []
        pts = clip_edge(pts, inside, inter)
    return pts


def _area(pts):
    if len(pts) < 3:
        return 0.0
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _brute_force_rows(O, im, src, lat):
    rows = []
    ox = [0, 1, 2, 2, 2, 1, 0, 0, 1]
    oy = [0, 0, 0, 1, 2, 2, 2, 1, 1]
    for i in range(im.nx):
        for j in range(im.ny):
            sp = [lat[2 * i + ox[k], 2 * j + oy[k]] for k in range(9)]
            acc = {}
            for t in range(8):
                tri = [tuple(sp[(t + 1) % 8]), tuple(sp[t]), tuple(sp[8])]
                at = _area(tri)
                if at == 0:
                    continue
                for a in range(src.nx):
                    for b in range(src.ny):
                        x0 = src.xmin + a * src.dx
                        y0 = src.ymin + b * src.dy
                        c = _area(_clip(tri, x0, x0 + src.dx, y0, y0 + src.dy))
                        if c > 0:
                            acc[src.ny * a + b] = acc.get(src.ny * a + b, 0.0) + 0.25 * c / at
            rows.append(acc)
    return rows


@pytest.mark.parametrize('case', ['identity_shifted', 'lensed'])
def test_geometry_matches_polygon_clipping(oracle, case):
    O = oracle
    im = O.Grid2D(6, 7, [-0.6, 0.6, -0.7, 0.7])
    src = O.Grid2D(7, 6, [-0.613, 0.587, -0.5871, 0.6129])
    lat = O.stencil_lattice(im)
    if case == 'lensed':
        L = [O.pack_pemd(b=0.35, q=0.8, th=25.0, qh=0.5, x=0.031, y=-0.017)]
        lat = O.deflect_points(L, lat.reshape(-1, 2)).reshape(lat.shape)
    rp, ci, va, st = O.lens_matrix_from_lattice_nodrop(im, src, lat)
    assert st == 0
    rows = _brute_force_rows(O, im, src, lat)
    for r, ref in enumerate(rows):
        cols = ci[rp[r]:rp[r + 1]]
        vals = va[rp[r]:rp[r + 1]]
        got = dict(zip(cols.tolist(), vals.tolist()))
        if not ref:
            continue
        for c, v in ref.items():
            assert c in got, (r, c)
            assert abs(got[c] - v) < 1e-12, (r, c, got[c], v)
        for c, v in got.items():
            if c not in ref:
                assert abs(v) < 1e-12, (r, c, v)


def test_identity_interior_rows_sum_to_two(oracle):
    O = oracle
    im = O.Grid2D(10, 10, [-1, 1, -1, 1])
    src = O.Grid2D(13, 11, [-1.3071, 1.2929, -1.1013, 1.0987])
    lat = O.stencil_lattice(im)
    rp, ci, va, _ = O.lens_matrix_from_lattice_nodrop(im, src, lat)
    sums = np.add.reduceat(va, rp[:-1])
    assert np.allclose(sums, 2.0, rtol=0, atol=1e-13)


def test_drop_rule_removes_one_fragment_per_later_group(oracle):
    """Dropped-fragment quirk (raster.c:378-384): the dropped mass per row is
    the difference between the no-drop and the reference reduction; it is
    zero for the first (lowest) column and non-negative elsewhere."""
    O = oracle
    im = O.Grid2D(16, 16, [-0.8, 0.8, -0.8, 0.8])
    src = O.Grid2D(12, 12, [-0.41, 0.43, -0.39, 0.45])
    L = [O.pack_pemd(b=0.5, q=0.85, th=15.0, qh=0.5, x=0.013, y=0.021)]
    lat = O.deflect_points(L, O.stencil_lattice(im).reshape(-1, 2)).reshape(33, 33, 2)
    a = O.lens_matrix_from_lattice(im, src, lat)
    b = O.lens_matrix_from_lattice_nodrop(im, src, lat)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    diff = b[2] - a[2]
    for r in range(im.nx * im.ny):
        s, e = a[0][r], a[0][r + 1]
        if e > s:
            assert diff[s] == 0.0
            assert np.all(diff[s:e] >= -1e-15)


def _ref_or_skip(O):
    if O.ref_lib() is None:
        pytest.skip('oracle/_ref not built (reference sources absent)')


def test_oracle_matches_reference_build_c1(oracle):
    O = oracle
    _ref_or_skip(O)
    im = O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28])
    src = O.Grid2D(64, 64, [-0.64, 0.64, -0.64, 0.64])
    rp, ci, va, st = O.lens_matrix_csr(im, src, c1_lens(O))
    rr, cr, vr = O.ref_lens_matrix_csr(im, src, c1_lens(O))
    assert np.array_equal(rp, rr) and np.array_equal(ci, cr)
    assert np.array_equal(va.view(np.int64), vr.view(np.int64))
    assert np.diff(rp).max() > 1000  # the lens-centre pixel (long row) is exercised


def test_oracle_matches_reference_build_masked_and_empty_rows(oracle):
    """Masked rows have no entries; unmasked rows mapping outside the source
    grid keep one stale entry (raster.c:376,389-391)."""
    O = oracle
    _ref_or_skip(O)
    g = rng(21)
    mask = g.uniform(size=(40, 40)) > 0.3
    im = O.Grid2D(40, 40, [-1.0, 1.0, -1.0, 1.0], mask=mask)
    src = O.Grid2D(20, 20, [-0.2, 0.3, -0.25, 0.25])
    L = [O.pack_pemd(b=0.6, q=0.7, th=-40.0, qh=0.47, x=0.05, y=-0.02), O.pack_shear(ss=0.02, sa=10.0)]
    rp, ci, va, st = O.lens_matrix_csr(im, src, L)
    rr, cr, vr = O.ref_lens_matrix_csr(im, src, L)
    assert np.array_equal(rp, rr) and np.array_equal(ci, cr)
    assert np.array_equal(va.view(np.int64), vr.view(np.int64))
    nn = np.diff(rp)
    assert np.all(nn[~mask.reshape(-1)] == 0)
    assert np.any((nn == 1) & (np.add.reduceat(np.abs(va), rp[:-1]) == 0) & mask.reshape(-1))
