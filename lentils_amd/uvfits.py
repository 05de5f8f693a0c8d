"""uvfits (random-groups FITS) reader for radio visibility data.

Mirrors the reference's loader (lentils/common/data_util.py:6-76), which reads
the file with astropy.io.fits: the primary HDU holds GCOUNT groups of PCOUNT
random parameters (UU, VV, WW in seconds, ...) followed by the data array
(COMPLEX x STOKES x FREQ x IF x RA x DEC, each COMPLEX entry = re, im,
weight); the 'AIPS FQ' binary table gives the IF frequency offsets and channel
widths.  astropy is not part of this environment, so the FITS structures are
parsed here directly (big-endian, 2880-byte records).

Returns the same four objects as the reference: a VisibilitySpace (channel
frequencies in Hz, uvw in seconds, so that u = U * frequency is in
wavelengths), the visibilities [channel][stokes][row] complex128, sigma =
weight^-1/2 and the mask weight > 0, with RR and LL combined as
0.5 (RR + LL), sigma 0.5 sqrt(s_RR^2 + s_LL^2) when combine_stokes is set.
"""
import numpy as np

from .fitsio import read_fits_header
from .spaces import VisibilitySpace

_DTYPES = {8: '>u1', 16: '>i2', 32: '>i4', 64: '>i8', -32: '>f4', -64: '>f8'}
# binary-table column codes: (numpy dtype, bytes)
_TFORM = {'L': ('>u1', 1), 'B': ('>u1', 1), 'I': ('>i2', 2), 'J': ('>i4', 4), 'K': ('>i8', 8), 'A': ('S1', 1),
          'E': ('>f4', 4), 'D': ('>f8', 8), 'C': ('>c8', 8), 'M': ('>c16', 16)}


def _padded(nbytes):
    return (nbytes + 2879) // 2880 * 2880


def _hdu_data_bytes(h):
    naxis = int(h.get('NAXIS', 0))
    if naxis == 0:
        return 0
    dims = [int(h[f'NAXIS{i}']) for i in range(1, naxis + 1)]
    bits = abs(int(h['BITPIX']))
    if h.get('GROUPS') and dims[0] == 0:  # random groups: NAXIS1 = 0
        n = int(np.prod(dims[1:])) if len(dims) > 1 else 1
        return (int(h.get('PCOUNT', 0)) + n) * int(h.get('GCOUNT', 1)) * bits // 8
    return (int(np.prod(dims)) * bits // 8 + int(h.get('PCOUNT', 0))) * int(h.get('GCOUNT', 1))


def _parse_tform(t):
    t = t.strip()
    k = 0
    while k < len(t) and t[k].isdigit():
        k += 1
    rep = int(t[:k]) if k else 1
    code = t[k]
    if code not in _TFORM:
        raise ValueError(f'unsupported binary-table column format {t!r}')
    return rep, code


def _read_bintable(f, h):
    """Columns of a BINTABLE HDU whose header has just been read: name -> array."""
    rowlen, nrows = int(h['NAXIS1']), int(h['NAXIS2'])
    raw = f.read(_padded(_hdu_data_bytes(h)))
    cols, off = {}, 0
    for c in range(1, int(h['TFIELDS']) + 1):
        rep, code = _parse_tform(h[f'TFORM{c}'])
        dt, nb = _TFORM[code]
        name = str(h.get(f'TTYPE{c}', f'COL{c}')).strip()
        width = rep * nb if code != 'A' else rep
        a = np.empty((nrows, rep) if code != 'A' else (nrows,), dtype=dt if code != 'A' else f'S{rep}')
        for r in range(nrows):
            chunk = raw[r * rowlen + off: r * rowlen + off + width]
            if code == 'A':
                a[r] = chunk
            else:
                a[r] = np.frombuffer(chunk, dtype=dt, count=rep)
        cols[name] = a
        off += width
    return cols


def read_uvfits(path):
    """Raw content: (primary header, random parameters {name: float64 [G]},
    data float64 [G][DEC][RA][IF][FREQ][STOKES][COMPLEX], {extname: columns})."""
    with open(path, 'rb') as f:
        h = read_fits_header(f)
        if not h.get('GROUPS') or int(h.get('NAXIS1', -1)) != 0:
            raise ValueError(f'{path}: not a random-groups (uvfits) file')
        naxis = int(h['NAXIS'])
        dims = [int(h[f'NAXIS{i}']) for i in range(naxis, 1, -1)]  # slowest first
        pcount, gcount = int(h['PCOUNT']), int(h['GCOUNT'])
        dt = np.dtype(_DTYPES[int(h['BITPIX'])])
        ndata = int(np.prod(dims))
        raw = f.read(_padded(_hdu_data_bytes(h)))
        grp = np.frombuffer(raw[:(pcount + ndata) * gcount * dt.itemsize], dtype=dt).reshape(gcount, pcount + ndata)
        params = {}
        for k in range(pcount):
            name = str(h[f'PTYPE{k + 1}']).strip()
            v = grp[:, k].astype(np.float64)
            sc, z = float(h.get(f'PSCAL{k + 1}', 1.0)), float(h.get(f'PZERO{k + 1}', 0.0))
            if sc != 1.0 or z != 0.0:
                v = v * sc + z
            params[name] = params[name] + v if name in params else v  # repeated names add (e.g. DATE)
        data = grp[:, pcount:].astype(np.float64).reshape([gcount] + dims)
        bs, bz = float(h.get('BSCALE', 1.0)), float(h.get('BZERO', 0.0))
        if bs != 1.0 or bz != 0.0:
            data = data * bs + bz
        tables = {}
        while True:
            try:
                xh = read_fits_header(f)
            except ValueError:
                break
            if xh.get('XTENSION', '').strip() == 'BINTABLE':
                tables[str(xh.get('EXTNAME', '')).strip()] = _read_bintable(f, xh)
            else:
                f.seek(_padded(_hdu_data_bytes(xh)), 1)
    return h, params, data, tables


def load_uvfits(path, combine_stokes=True):
    """(VisibilitySpace, data, sigma, mask) as data_util.py:6-76 returns them;
    data / sigma / mask are [channel][stokes][row]."""
    h, params, data, tables = read_uvfits(path)
    axes = {}
    for ax in range(1, int(h['NAXIS']) + 1):
        if int(h[f'NAXIS{ax}']) == 0:
            continue
        axes[str(h[f'CTYPE{ax}']).strip()] = {'NAXIS': int(h[f'NAXIS{ax}']), 'CRVAL': float(h[f'CRVAL{ax}']),
                                              'CRPIX': float(h[f'CRPIX{ax}'])}
    num_rows = int(h['GCOUNT'])
    num_spw = axes['IF']['NAXIS']
    num_channels = axes['FREQ']['NAXIS']
    ref_freq = axes['FREQ']['CRVAL']
    ref_ch = axes['FREQ']['CRPIX'] - 1
    fq = tables['AIPS FQ']
    chw = np.asarray(fq['CH WIDTH'], dtype=np.float64).reshape(num_spw)
    iff = np.asarray(fq['IF FREQ'], dtype=np.float64).reshape(num_spw, 1)
    channels = (np.multiply.outer(chw, np.arange(num_channels) - ref_ch) + iff + ref_freq).flatten()
    num_channels *= num_spw
    uvcoords = np.array([params['UU'], params['VV'], params['WW']]).T.astype(np.float64, order='C')
    d = data[:, 0, 0]  # [row][IF][FREQ][STOKES][COMPLEX]
    vis = d[..., 0] + 1j * d[..., 1]
    w = d[..., 2]
    mask = w > 0.0
    sigma = np.zeros_like(w)
    sigma[mask] = w[mask] ** -0.5
    if combine_stokes:
        mask = np.all(mask, axis=-1)
        vis = 0.5 * (vis[..., 0] + vis[..., 1])
        s_rr, s_ll = sigma[..., 0], sigma[..., 1]
        sigma = np.zeros_like(s_rr)
        sigma[mask] = 0.5 * np.sqrt(s_rr[mask] ** 2 + s_ll[mask] ** 2)
    else:
        vis, sigma, mask = vis[..., 0], sigma[..., 0], mask[..., 0]
    space = VisibilitySpace(channels=channels, uvcoords=uvcoords)

    def order(a, dtype):
        return np.ascontiguousarray(np.moveaxis(a.reshape(num_rows, space.num_channels, space.num_stokes),
                                                [0, 1, 2], [2, 0, 1]), dtype=dtype)
    return space, order(vis, np.complex128), order(sigma, np.float64), order(mask, np.bool_)
