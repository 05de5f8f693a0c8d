"""Minimal FITS reader for 2-D/3-D primary-HDU images (masks, dirty images).

astropy is not part of this environment; the reference reads masks with
astropy.io.fits (space_base.py:109-111).  This reader handles the simple
primary HDU case those files use: 2880-byte header cards, BITPIX
8/16/32/64/-32/-64, big-endian data, BSCALE/BZERO.
"""
import numpy as np

_DTYPES = {8: '>u1', 16: '>i2', 32: '>i4', 64: '>i8', -32: '>f4', -64: '>f8'}


def read_fits_header(f):
    cards = {}
    while True:
        block = f.read(2880)
        if len(block) < 2880:
            raise ValueError('truncated FITS header')
        for k in range(36):
            card = block[80 * k:80 * (k + 1)].decode('ascii', 'replace')
            key = card[:8].strip()
            if key == 'END':
                return cards
            if card[8:10] == '= ':
                val = card[10:].split('/')[0].strip()
                if val.startswith("'"):
                    cards[key] = val.strip("'").strip()
                elif val in ('T', 'F'):
                    cards[key] = val == 'T'
                else:
                    try:
                        cards[key] = int(val)
                    except ValueError:
                        try:
                            cards[key] = float(val)
                        except ValueError:
                            cards[key] = val


def read_fits_image(path):
    """Primary-HDU data as a numpy array in FITS (C, slowest-first) order."""
    with open(path, 'rb') as f:
        h = read_fits_header(f)
        naxis = int(h.get('NAXIS', 0))
        dims = [int(h[f'NAXIS{i}']) for i in range(naxis, 0, -1)]
        dt = np.dtype(_DTYPES[int(h['BITPIX'])])
        n = int(np.prod(dims)) if dims else 0
        data = np.frombuffer(f.read(n * dt.itemsize), dtype=dt).reshape(dims)
    out = data.astype(np.float64 if dt.kind == 'f' else dt.newbyteorder('='))
    if 'BSCALE' in h or 'BZERO' in h:
        out = out * float(h.get('BSCALE', 1.0)) + float(h.get('BZERO', 0.0))
    return out
