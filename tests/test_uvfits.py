"""uvfits loading (data_util.py:6-76): the random-groups reader against a
file written here with known content, and the reference's two test files
(tests/golden/*.uvfits, copied from lentils/tests/data_radio_2d) for their
documented facts; on the GPU, the real uv set through the NUFFT against the
oracle."""
import os

import numpy as np
import pytest

from lentils_amd.uvfits import load_uvfits, read_uvfits

HERE = os.path.dirname(os.path.abspath(__file__))
TINY = os.path.join(HERE, 'golden', 'j0751_lwmp_nopos_nosub_0.000000_tiny.uvfits')
SMALL = os.path.join(HERE, 'golden', 'j0751_small_snr1x.uvfits')


def _card(key, val):
    if isinstance(val, bool):
        v = ('T' if val else 'F').rjust(20)
    elif isinstance(val, str):
        v = f"'{val:<8}'"
    elif isinstance(val, int):
        v = str(val).rjust(20)
    else:
        v = repr(float(val)).upper().rjust(20)
    return f'{key:<8}= {v}'.ljust(80)[:80].encode('ascii')


def _header(cards):
    b = b''.join(_card(k, v) for k, v in cards) + 'END'.ljust(80).encode('ascii')
    return b + b' ' * ((-len(b)) % 2880)


def _pad(b):
    return b + b'\0' * ((-len(b)) % 2880)


def _write_uvfits(path, uvw, vis, w, nif, nch, if_freq, ch_width, ref_freq, pscal_uu=1.0, pzero_uu=0.0):
    """Minimal uvfits writer: uvw [G][3] (seconds), vis / w [G][IF][FREQ][STOKES]."""
    G, nst = vis.shape[0], vis.shape[3]
    cards = [('SIMPLE', True), ('BITPIX', -32), ('NAXIS', 7), ('NAXIS1', 0), ('NAXIS2', 3), ('NAXIS3', nst),
             ('NAXIS4', nch), ('NAXIS5', nif), ('NAXIS6', 1), ('NAXIS7', 1), ('EXTEND', True), ('GROUPS', True),
             ('PCOUNT', 4), ('GCOUNT', G), ('BSCALE', 1.0), ('BZERO', 0.0)]
    for ax, (ct, cv, cp) in enumerate([('COMPLEX', 1.0, 1.0), ('STOKES', -1.0, 1.0), ('FREQ', ref_freq, 1.0),
                                       ('IF', 1.0, 1.0), ('RA', 0.0, 1.0), ('DEC', 0.0, 1.0)], start=2):
        cards += [(f'CTYPE{ax}', ct), (f'CRVAL{ax}', cv), (f'CDELT{ax}', 1.0), (f'CRPIX{ax}', cp),
                  (f'CROTA{ax}', 0.0)]
    for k, name in enumerate(['UU', 'VV', 'WW', 'DATE'], start=1):
        cards += [(f'PTYPE{k}', name), (f'PSCAL{k}', pscal_uu if k == 1 else 1.0),
                  (f'PZERO{k}', pzero_uu if k == 1 else 0.0)]
    par = np.zeros((G, 4), dtype='>f4')
    par[:, 0] = (uvw[:, 0] - pzero_uu) / pscal_uu
    par[:, 1:3] = uvw[:, 1:3]
    dat = np.zeros((G, 1, 1, nif, nch, nst, 3), dtype='>f4')
    dat[..., 0] = vis.real.reshape(G, 1, 1, nif, nch, nst)
    dat[..., 1] = vis.imag.reshape(G, 1, 1, nif, nch, nst)
    dat[..., 2] = w.reshape(G, 1, 1, nif, nch, nst)
    body = np.concatenate([par, dat.reshape(G, -1)], axis=1).astype('>f4').tobytes()
    fq_cards = [('XTENSION', 'BINTABLE'), ('BITPIX', 8), ('NAXIS', 2), ('NAXIS1', 4 + 12 * nif), ('NAXIS2', 1),
                ('PCOUNT', 0), ('GCOUNT', 1), ('TFIELDS', 3), ('TTYPE1', 'FRQSEL'), ('TFORM1', '1J'),
                ('TTYPE2', 'IF FREQ'), ('TFORM2', f'{nif}D'), ('TTYPE3', 'CH WIDTH'), ('TFORM3', f'{nif}E'),
                ('EXTNAME', 'AIPS FQ')]
    row = (np.array([1], dtype='>i4').tobytes() + np.asarray(if_freq, dtype='>f8').tobytes()
           + np.asarray(ch_width, dtype='>f4').tobytes())
    with open(path, 'wb') as f:
        f.write(_header(cards))
        f.write(_pad(body))
        f.write(_header(fq_cards))
        f.write(_pad(row))


def test_reader_round_trip(tmp_path):
    g = np.random.default_rng(3)
    G, nif, nch = 57, 2, 3
    uvw = g.normal(0, 1e-4, size=(G, 3)).astype(np.float32).astype(np.float64)
    vis = (g.normal(size=(G, nif, nch, 2)) + 1j * g.normal(size=(G, nif, nch, 2))).astype(np.complex64)
    w = g.uniform(0.5, 4.0, size=(G, nif, nch, 2)).astype(np.float32)
    w[::7, 0, 1, 1] = -1.0  # flagged
    path = str(tmp_path / 't.uvfits')
    _write_uvfits(path, uvw, vis, w, nif, nch, [0.0, 8e6], [1e6, 1e6], 1.5e9, pscal_uu=0.5, pzero_uu=0.0)
    h, params, data, tables = read_uvfits(path)
    assert data.shape == (G, 1, 1, nif, nch, 2, 3)
    assert np.allclose(params['UU'], uvw[:, 0], rtol=1e-7, atol=0)
    space, d, sig, m = load_uvfits(path, combine_stokes=True)
    assert space.num_channels == nif * nch and space.num_rows == G
    want_ch = (np.multiply.outer(np.array([1e6, 1e6]), np.arange(nch)) + np.array([[0.0], [8e6]]) + 1.5e9).ravel()
    assert np.array_equal(space.channels, want_ch)
    v = vis.astype(np.complex128)
    rr, ll = v[..., 0], v[..., 1]
    want = np.moveaxis((0.5 * (rr + ll)).reshape(G, nif * nch, 1), [0, 1, 2], [2, 0, 1])
    assert np.allclose(d, want, rtol=1e-6, atol=0)
    mk = np.all(w > 0, axis=-1).reshape(G, nif * nch)
    assert np.array_equal(m[:, 0, :], mk.T)
    wpos = np.where(w > 0, w, 1.0).astype(np.float64)
    s2 = 0.5 * np.sqrt(1.0 / wpos[..., 0] + 1.0 / wpos[..., 1]).reshape(G, -1)
    assert np.allclose(sig[:, 0, :][mk.T], s2.T[mk.T], rtol=1e-6)
    assert np.all(sig[:, 0, :][~mk.T] == 0)
    # one Stokes parameter only
    space1, d1, sig1, m1 = load_uvfits(path, combine_stokes=False)
    assert np.allclose(d1[:, 0, :], v[..., 0].reshape(G, -1).T, rtol=1e-6, atol=0)


@pytest.mark.parametrize('path,rows', [(TINY, 1298), (SMALL, 3063)])
def test_reference_files(path, rows):
    """The j0751 test files: GCOUNT rows, two Stokes parameters, one IF of one
    channel at 1.63074 GHz (the header), u up to ~6e7 wavelengths."""
    h, params, data, tables = read_uvfits(path)
    assert data.shape == (rows, 1, 1, 1, 1, 2, 3)
    assert 'AIPS FQ' in tables
    space, d, sig, m = load_uvfits(path)
    assert space.num_rows == rows and space.num_channels == 1
    assert np.isclose(space.channels[0], 1.63074e9)
    w = data[:, 0, 0, 0, 0, :, 2]
    assert np.array_equal(m[0, 0], np.all(w > 0, axis=-1))
    assert 0.9 < m.mean() <= 1.0
    uvmax = np.abs(space.uvw[:, :2]).max() * space.channels[0]
    assert 1e7 < uvmax < 1e8
    assert np.all(np.isfinite(d)) and np.all(sig[m] > 0)


@pytest.mark.gpu
def test_real_uv_set_through_nufft(hip, oracle):
    """The tiny file's uv set through the FP64 NUFFT plan: the dirty image of
    its (mask-weighted) data equals the oracle's to 1e-12."""
    from lentils_amd import ImageSpace
    from lentils_amd.solvers import NUFFTPlan
    O = oracle
    space, d, sig, m = load_uvfits(TINY)
    # 128^2 image at 1.2 mas (the reference test's pixel scale, test_all.py:30)
    im = ImageSpace(shape=(128, 128), bounds=[(-0.0768, 0.0768), (-0.0768, 0.0768)])
    plan = NUFFTPlan(space, im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f64')
    w = np.where(m, 1.0 / np.where(m, sig, 1.0) ** 2, 0.0).reshape(-1)
    dirty = plan.adjoint(w * d.reshape(-1))
    og = O.Grid2D(128, 128, [-0.0768, 0.0768, -0.0768, 0.0768])
    geo = O.NUFFTGeometry(og, pad=2, wsup=3, beta=14.04)
    ref = O.nufft_adjoint((w * d.reshape(-1)).reshape(1, -1), space.uvw, space.channels, geo, nthreads=8)
    assert O.max_relative_error(dirty.reshape(-1), ref.reshape(-1)) < 1e-12
