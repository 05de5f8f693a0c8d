"""Pins the oracle's deflection (CPU, no GPU).

* bit-for-bit against the reference's own deflect.c built by oracle/Makefile
  (when the reference sources were present at build time);
* against closed forms the mathematics fixes: the 2F1 series equals
  scipy.special.hyp2f1(1, 2-t, 2-t/2, u) up to its 20-term truncation,
  a circular power law has |alpha| = b_eff^t r^(1-t), the isothermal
  circular case |alpha| = b, external shear is linear, rotating the lens and
  the point together rotates the deflection.
"""
import numpy as np
import pytest
import scipy.special as sps

from helpers import c1_lens, random_lenses, rng


def test_oracle_matches_reference_build_bitwise(oracle):
    O = oracle
    if O.ref_lib() is None:
        pytest.skip('oracle/_ref not built (reference sources absent)')
    g = rng(1)
    pts = g.uniform(-1.3, 1.3, size=(20000, 2))
    cases = [c1_lens(O)] + [[O.pack_pemd(b=1.2, q=q, th=77.0, qh=qh, x=-0.03, y=0.04),
                             O.pack_shear(ss=0.04, sa=12.0, x=-0.03, y=0.04, gks=0.01, gka=20.0, gss=0.02,
                                          gsa=-30.0)]
                            for q, qh in [(0.5, 0.4), (0.95, 0.6), (0.3, 0.55), (0.99, 0.5)]]
    for lenses in cases:
        a = O.deflect_points(lenses, pts)
        b = O.deflect_points(lenses, pts, use_ref=True)
        assert np.array_equal(a.view(np.int64), b.view(np.int64))


def _alpha(O, lenses, pts):
    return pts - O.deflect_points(lenses, pts)


def test_sis_circular_closed_form(oracle):
    O = oracle
    # q = 1, gamma = 2 (t = 1): |alpha| = b everywhere (singular isothermal sphere)
    L = [O.pack_pemd(b=0.7, q=1.0, th=0.0, qh=0.5, x=0.1, y=-0.2)]
    pts = rng(2).uniform(-1, 1, size=(1000, 2))
    a = _alpha(O, L, pts)
    assert np.allclose(np.hypot(a[:, 0], a[:, 1]), 0.7, rtol=1e-14, atol=0)
    # radial direction
    d = pts - np.array([0.1, -0.2])
    cross = a[:, 0] * d[:, 1] - a[:, 1] * d[:, 0]
    assert np.max(np.abs(cross)) < 1e-13


@pytest.mark.parametrize('qh', [0.4, 0.55, 0.6])
def test_circular_power_law(oracle, qh):
    O = oracle
    t = 2 * qh
    b = 0.9
    L = [O.pack_pemd(b=b, q=1.0, th=10.0, qh=qh, x=0.0, y=0.0)]
    pts = rng(3).uniform(-1, 1, size=(500, 2))
    a = _alpha(O, L, pts)
    beff = (0.5 * b * (3.0 - t) / (2.0 - t)) ** (1.0 / t)
    r = np.hypot(pts[:, 0], pts[:, 1])
    assert np.allclose(np.hypot(a[:, 0], a[:, 1]), beff ** t * r ** (1 - t), rtol=1e-13)


def test_series_against_library_hyp2f1(oracle):
    """deflect.c:179-239 sums 20 terms of 2F1(1, 2-t; 2-t/2; u); for q >= 0.7
    the truncation is below 1e-15, so the PEMD deflection built from the
    library 2F1 must agree to ~1e-13."""
    O = oracle
    g = rng(4)
    for q, qh in [(0.8, 0.5), (0.9, 0.45), (0.75, 0.6)]:
        t = 2 * qh
        b = 1.1
        L = [O.pack_pemd(b=b, q=q, th=0.0, qh=qh, x=0.0, y=0.0)]
        pts = g.uniform(-1, 1, size=(200, 2))
        a = _alpha(O, L, pts)
        # th = 0: crot = cos(90) - i sin(90) = -i; z = (x + i y)(-i)
        beff = (0.5 * b * q ** (t - 0.5) * (3.0 - t) / (2.0 - t)) ** (1.0 / t)
        z = (pts[:, 0] + 1j * pts[:, 1]) * (np.cos(np.pi / 2) - 1j * np.sin(np.pi / 2))
        x, y = z.real, z.imag
        rs = np.sqrt(q * q * x * x + y * y)
        w2 = (1 - q * q) / q ** 2 * rs ** 2 / z ** 2
        u = 0.5 * (1 - np.sqrt(1 - w2))
        F = np.array([complex(sps.hyp2f1(1, 2 - t, 2 - t / 2, ui)) if abs(ui.imag) < 1e-300
                      else _hyp2f1_complex(t, ui) for ui in u])
        A = beff ** 2 / (q * z) * (beff / rs) ** (t - 2)
        alpha = np.conj(A * F * (np.cos(np.pi / 2) - 1j * np.sin(np.pi / 2)))
        ref = np.stack([alpha.real, alpha.imag], 1)
        assert np.max(np.abs(a - ref)) / np.max(np.abs(ref)) < 1e-12


def _hyp2f1_complex(t, u):
    # scipy.special.hyp2f1 accepts complex z
    return complex(sps.hyp2f1(1.0, 2.0 - t, 2.0 - t / 2.0, complex(u)))


def test_shear_linear_and_rotation_covariance(oracle):
    O = oracle
    g = rng(6)
    pts = g.uniform(-1, 1, size=(300, 2))
    S = [O.pack_shear(ss=0.05, sa=30.0, x=0.1, y=0.2)]
    a = _alpha(O, S, pts)
    c, s = np.cos(2 * np.deg2rad(30.0)), np.sin(2 * np.deg2rad(30.0))
    d = pts - np.array([0.1, 0.2])
    ref = 0.05 * np.stack([c * d[:, 0] + s * d[:, 1], s * d[:, 0] - c * d[:, 1]], 1)
    assert np.allclose(a, ref, rtol=1e-12, atol=1e-15)
    # rotating lens and point by 20 degrees rotates the deflection
    L0 = [O.pack_pemd(b=1.0, q=0.8, th=10.0, qh=0.5, x=0.0, y=0.0)]
    L1 = [O.pack_pemd(b=1.0, q=0.8, th=30.0, qh=0.5, x=0.0, y=0.0)]
    ang = np.deg2rad(20.0)
    R = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    a0 = _alpha(O, L0, pts)
    a1 = _alpha(O, L1, pts @ R.T)
    assert np.max(np.abs(a1 - a0 @ R.T)) < 1e-12


def test_components_add_linearly(oracle):
    O = oracle
    pts = rng(7).uniform(-1, 1, size=(300, 2))
    P = O.pack_pemd(b=1.0, q=0.7, th=40.0, qh=0.52, x=0.01, y=0.02)
    S = O.pack_shear(ss=0.03, sa=45.0, x=0.01, y=0.02)
    a = _alpha(O, [P, S], pts)
    assert np.allclose(a, _alpha(O, [P], pts) + _alpha(O, [S], pts), rtol=1e-13, atol=1e-15)


def _alpha(O, lenses, pts):
    return pts - O.deflect_points(lenses, pts)


def test_deflection_derivatives_match_reference_build(oracle):
    """want_deriv = 1 (deflect.c:179-239, 243-338): the oracle's sums equal
    the reference's own deflect.c (oracle/_ref) bit for bit, and so do the
    deflected points."""
    O = oracle
    if O.ref_lib() is None:
        pytest.skip('reference sources absent: oracle/_ref not built')
    for k, lenses in enumerate([c1_lens(O)] + random_lenses(O, 4)):
        pts = rng(50 + k).uniform(-2.2, 2.2, size=(4000, 2))
        a, d, per = O.deflect_points_deriv(lenses, pts)
        r, dr, _ = O.deflect_points_deriv(lenses, pts, use_ref=True)
        assert np.array_equal(a, r)
        assert np.array_equal(d, dr)
        assert np.all(np.isfinite(per))
        assert np.array_equal(a, O.deflect_points(lenses, pts))


def _pemd_beff(O, beff, q, th, qh, x, y):
    """A PEMD record whose internal radius (deflect.c:284) is beff: the
    reference differentiates in (b_eff, qh, q, x, y, th) with b_eff held
    fixed (dA_db = A t / b, deflect.c:316)."""
    t = 2 * qh
    b = 2 * beff ** t * (2 - t) / (q ** (t - 0.5) * (3 - t))
    return O.pack_pemd(b=b, q=q, th=th, qh=qh, x=x, y=y)


def test_deflection_derivatives_finite_differences(oracle):
    """Each point's d alpha / d(b_eff, qh, q, x, y, th) against central
    finite differences of the deflection (th in degrees, qh the slope with
    t = 2 qh, b_eff the internal radius held fixed while q and qh move): a
    dropped term, a wrong sign or a swapped component fails this."""
    O = oracle
    g = rng(60)
    for case in range(3):
        beff, q, th, qh = g.uniform(0.6, 1.4), g.uniform(0.6, 0.95), g.uniform(0, 180), g.uniform(0.4, 0.6)
        x0, y0 = g.uniform(-0.05, 0.05, 2)
        shear = O.pack_shear(ss=0.03, sa=g.uniform(0, 180), x=x0, y=y0)
        pts = g.uniform(-2.0, 2.0, size=(600, 2))
        pts = pts[np.hypot(pts[:, 0] - x0, pts[:, 1] - y0) > 0.1]
        base = dict(beff=beff, q=q, th=th, qh=qh, x=x0, y=y0)
        _, _, per = O.deflect_points_deriv([_pemd_beff(O, **base), shear], pts)
        steps = {'beff': 1e-6, 'qh': 1e-6, 'q': 1e-6, 'x': 1e-6, 'y': 1e-6, 'th': 1e-4}
        for k, name in enumerate(['beff', 'qh', 'q', 'x', 'y', 'th']):
            h = steps[name]
            up, dn = dict(base), dict(base)
            up[name] += h
            dn[name] -= h
            fd = (_alpha(O, [_pemd_beff(O, **up), shear], pts) -
                  _alpha(O, [_pemd_beff(O, **dn), shear], pts)) / (2 * h)
            an = per[:, 2 * k:2 * k + 2]
            err = np.max(np.abs(an - fd)) / np.max(np.abs(fd))
            assert err < 1e-6, (case, name, err)
