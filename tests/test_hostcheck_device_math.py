"""CPU check of the device arithmetic: deflect_math.h and raster_core.h (the
code the HIP kernels execute) compiled for the host and compared bit for bit
with the oracle.  On the GPU, pow/hypot come from the device math library
instead of glibc, so the GPU tests use a 1e-12 relative bar for
deflections; the rasterizer has no libm calls besides floor, so its host
build must match the oracle exactly."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import c1_lens, random_lenses, rng

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'native', 'hostcheck.hip')
OUT = os.path.join(HERE, 'native', '_build', 'libhostcheck.so')


@pytest.fixture(scope='module')
def hc():
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(hipcc):
        pytest.skip('hipcc not available')
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    deps = [SRC] + [os.path.join(HERE, '..', 'lentils_amd', 'csrc', f) for f in ('deflect_math.h', 'raster_core.h',
                                                                                 'lt_internal.h')]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        cmd = [hipcc, '-O2', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off', '-fwrapv',
               '--offload-arch=gfx950', SRC, '-o', OUT]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    L = C.CDLL(OUT)
    L.hc_deflect_points.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_void_p]
    L.hc_deflect_points.restype = C.c_int
    L.hc_deflect_points_deriv.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
    L.hc_deflect_points_deriv.restype = C.c_int
    L.hc_lens_matrix_from_lattice.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.hc_lens_matrix_from_lattice.restype = C.c_long
    L.hc_pixel_frags.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_long,
                                 C.POINTER(C.c_int)]
    L.hc_pixel_frags.restype = C.c_long
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_device_deflection_math_bitwise_on_host(hc, oracle):
    O = oracle
    pts = rng(31).uniform(-1.5, 1.5, size=(20000, 2))
    for lenses in [c1_lens(O)] + random_lenses(O, 6):
        arr = np.array(lenses, dtype=O.MASS_MODEL_DTYPE)
        out = np.zeros_like(pts)
        assert hc.hc_deflect_points(_p(arr), len(arr), _p(pts), pts.shape[0], _p(out)) == 0
        ref = O.deflect_points(lenses, pts)
        assert np.array_equal(out.view(np.int64), ref.view(np.int64))


def test_device_raster_code_bitwise_on_host(hc, oracle):
    O = oracle
    im = O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28])
    src = O.Grid2D(64, 64, [-0.64, 0.64, -0.64, 0.64])
    lat = O.deflect_points(c1_lens(O), O.stencil_lattice(im).reshape(-1, 2)).reshape(257, 257, 2)
    rr, cr, vr, _ = O.lens_matrix_from_lattice(im, src, lat)
    rp = np.zeros(128 * 128 + 1, dtype=np.int32)
    st = C.c_int(0)
    mask = im.mask.astype(np.int8)
    n = hc.hc_lens_matrix_from_lattice(128, 128, _p(mask), 64, 64, src.xmin, src.ymin, src.dx, src.dy, _p(lat),
                                       _p(rp), None, None, C.byref(st))
    ci = np.zeros(n, dtype=np.int32)
    va = np.zeros(n)
    hc.hc_lens_matrix_from_lattice(128, 128, _p(mask), 64, 64, src.xmin, src.ymin, src.dx, src.dy, _p(lat), _p(rp),
                                   _p(ci), _p(va), C.byref(st))
    assert np.array_equal(rp, rr) and np.array_equal(ci, cr)
    assert np.array_equal(va.view(np.int64), vr.view(np.int64))


def _frags(hc, lat, ny, i, j, src, mode, front, cap=4_000_000):
    area = np.zeros(cap)
    idx = np.zeros(cap, dtype=np.int32)
    lev = C.c_int(0)
    n = hc.hc_pixel_frags(_p(lat), ny, i, j, src.nx, src.ny, src.xmin, src.ymin, src.dx, src.dy, mode, front,
                          _p(area), _p(idx), cap, C.byref(lev))
    assert n <= cap
    return (None if n < 0 else (area[:n].copy(), idx[:n].copy())), lev.value


@pytest.mark.parametrize('case', ['c1', 'wide'])
def test_long_row_decomposition_reproduces_emission_order(hc, oracle, case):
    """The breadth-first frontier + per-subtree depth-first decomposition the
    GPU uses for long rows emits exactly the reference's fragment sequence
    (same areas bit for bit, same order), for frontier caps from 1 to 512."""
    O = oracle
    if case == 'c1':
        im = O.Grid2D(128, 128, [-1.28, 1.28, -1.28, 1.28])
        src = O.Grid2D(64, 64, [-0.64, 0.64, -0.64, 0.64])
        lenses = c1_lens(O)
        ci, cj = 65, 63
    else:  # bench geometry around the lens centre: 0.01" pixels, 128^2 source
        im = O.Grid2D(96, 96, [-0.48, 0.48, -0.48, 0.48])
        src = O.Grid2D(128, 128, [-0.64, 0.64, -0.64, 0.64])
        lenses = random_lenses(O, 1, seed=5)[0]
        fp = np.array(lenses, dtype=O.MASS_MODEL_DTYPE)[0]['fpars']
        ci, cj = int((fp[0] - im.xmin) / im.dx), int((fp[1] - im.ymin) / im.dy)
    n = im.nx
    lat = O.deflect_points(lenses, O.stencil_lattice(im).reshape(-1, 2)).reshape(2 * n + 1, 2 * n + 1, 2)
    lat = np.ascontiguousarray(lat)
    g = rng(33)
    pix = [(ci + a, cj + b) for a in range(-2, 3) for b in range(-2, 3)]
    pix += [tuple(v) for v in g.integers(0, n, size=(10, 2))]
    longest = 0
    compared = 0
    for (i, j) in pix:
        ref, _ = _frags(hc, lat, n, i, j, src, 0, 0)
        longest = max(longest, ref[0].size)
        for front in (1, 4, 16, 512, 2048):
            got, lev = _frags(hc, lat, n, i, j, src, 1, front)
            if got is None:  # the GPU would take the sequential path for this row
                continue
            assert np.array_equal(got[1], ref[1]), (i, j, front)
            assert np.array_equal(got[0].view(np.int64), ref[0].view(np.int64)), (i, j, front)
            compared += ref[0].size > 1000
    assert longest > 1000 and compared >= 4


def test_device_derivative_math_bitwise_on_host(hc, oracle):
    """The kernels' derivative arithmetic (pemd_deflection_d, want_deriv
    branch of deflect.c:179-338) built for the host equals the oracle bit
    for bit, per point (and so the reference build's sums)."""
    O = oracle
    pts = rng(32).uniform(-1.8, 1.8, size=(20000, 2))
    for lenses in [c1_lens(O)] + random_lenses(O, 6):
        arr = np.array(lenses, dtype=O.MASS_MODEL_DTYPE)
        out = np.zeros_like(pts)
        dv = np.zeros((pts.shape[0], 12))
        assert hc.hc_deflect_points_deriv(_p(arr), len(arr), _p(pts), pts.shape[0], _p(out), _p(dv)) == 0
        ref, _, per = O.deflect_points_deriv(lenses, pts)
        assert np.array_equal(out, ref)
        assert np.array_equal(dv, per)
