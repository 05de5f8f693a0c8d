"""CPU oracle for the Lentils forward-model + inversion path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only to check or to time the CPU path.  The
product package never imports it.

Two layers:

* ctypes bindings to ``liboracle.so`` (``lentils_oracle.c``): deflection,
  rasterization into the CSR lens matrix, KB gridding, DFT / curvature /
  gradient CSR builders -- an FP64 restatement of ``lentils/backend/*.c``.
* numpy restatements of the Python operator glue the reference builds on top
  (``lentils/operators/fourier_operators.py:12-126``, apodization, padding,
  FFT conventions) and of the CG solve the reference's tests run
  (``lentils/tests/test_all.py:63-74``).

Optionally ``_ref/libref_lens.so`` -- the reference's own ``deflect.c`` and
``raster.c`` compiled unchanged by ``oracle/Makefile`` -- is bound too, so
tests can pin this restatement bit for bit.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ARCSEC = 4.8481368111e-6  # lentils/backend/common.h:7, space_base.py:72

# generic_mass_model, lentils/backend/load_backend.py:55-56 (320 bytes)
MASS_MODEL_DTYPE = np.dtype([('type', 'u4'), ('_pad', 'u4'), ('z_l', 'f8'), ('z_s', 'f8'), ('d_l', 'f8'),
                             ('d_s', 'f8'), ('d_ls', 'f8'), ('sigma_c', 'f8'), ('beta', 'f8'),
                             ('fpars', 'f8', 24), ('ipars', 'i4', 8), ('flags', 'b', 32)])


class CImageSpace(C.Structure):
    _fields_ = [('nx', C.c_int), ('ny', C.c_int), ('nch', C.c_int), ('nst', C.c_int),
                ('xmin', C.c_double), ('xmax', C.c_double), ('ymin', C.c_double), ('ymax', C.c_double),
                ('dx', C.c_double), ('dy', C.c_double), ('channels', C.c_void_p), ('mask', C.c_void_p)]


class CFourierSpace(C.Structure):
    _fields_ = [('nu', C.c_int), ('nv', C.c_int), ('half_nv', C.c_int), ('nch', C.c_int), ('nst', C.c_int),
                ('du', C.c_double), ('dv', C.c_double), ('gcx', C.c_double), ('gcy', C.c_double),
                ('channels', C.c_void_p)]


class CVisSpace(C.Structure):
    _fields_ = [('nrows', C.c_int), ('nch', C.c_int), ('nst', C.c_int), ('uv', C.c_void_p),
                ('channels', C.c_void_p)]


class CDelaunaySpace(C.Structure):
    _fields_ = [('num_points', C.c_int), ('num_tris', C.c_int), ('nch', C.c_int), ('nst', C.c_int),
                ('channels', C.c_void_p), ('points', C.c_void_p), ('triangles', C.c_void_p)]


class CLensModel(C.Structure):
    _fields_ = [('num_lenses', C.c_int), ('num_zplanes', C.c_int), ('lenses', C.c_void_p)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Grid2D:
    """Plain description of a regular 2-D pixel grid (ImageSpace geometry).

    ``bounds`` = (xmin, xmax, ymin, ymax); dx = (xmax-xmin)/nx as in
    ``space_base.py:100-104``; rows/cols are x-major (index = ny*i + j)."""

    def __init__(self, nx, ny, bounds, mask=None):
        self.nx, self.ny = int(nx), int(ny)
        self.bounds = np.array(bounds, dtype=np.float64).flatten()
        self.xmin, self.xmax, self.ymin, self.ymax = [float(b) for b in self.bounds]
        self.dx = (self.xmax - self.xmin) / self.nx
        self.dy = (self.ymax - self.ymin) / self.ny
        self.mask = np.ones((self.nx, self.ny), dtype=np.bool_) if mask is None else \
            np.ascontiguousarray(mask, dtype=np.bool_)
        self._ch = np.zeros(1)

    def cstruct(self):
        return CImageSpace(self.nx, self.ny, 1, 1, self.xmin, self.xmax, self.ymin, self.ymax, self.dx, self.dy,
                           _ptr(self._ch), _ptr(self.mask))


def _load(path):
    lib = C.CDLL(path)
    return lib


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: run make -C oracle')
        _lib = _load(path)
        L = _lib
        L.oracle_deflect_points.argtypes = [CLensModel, C.c_void_p, C.c_int, C.c_void_p]
        L.oracle_deflect_points.restype = C.c_int
        L.oracle_deflect_points_deriv.argtypes = [CLensModel, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]
        L.oracle_deflect_points_deriv.restype = C.c_int
        L.oracle_triangle_gradient_csr.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_void_p]
        L.oracle_delaunay_lens_matrix_csr.argtypes = [CImageSpace, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_triangle_geometry_from_inds.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.oracle_lens_matrix_csr.argtypes = [CImageSpace, CImageSpace, CLensModel, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.POINTER(C.c_long)]
        L.oracle_lens_matrix_csr.restype = C.c_int
        L.oracle_lens_matrix_from_lattice.argtypes = [CImageSpace, CImageSpace, C.c_void_p, C.c_void_p,
                                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_long)]
        L.oracle_lens_matrix_from_lattice.restype = C.c_int
        L.oracle_lens_matrix_from_lattice_nodrop.argtypes = L.oracle_lens_matrix_from_lattice.argtypes
        L.oracle_lens_matrix_from_lattice_nodrop.restype = C.c_int
        L.oracle_stencil_lattice.argtypes = [CImageSpace, C.c_void_p]
        L.oracle_bessel_i0.argtypes = [C.c_double]
        L.oracle_bessel_i0.restype = C.c_double
        L.oracle_kb_g.argtypes = [C.c_double] * 4
        L.oracle_kb_g.restype = C.c_double
        L.oracle_grid.argtypes = [CVisSpace, C.c_void_p, CFourierSpace, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                  C.c_int]
        L.oracle_dft_matrix_csr.argtypes = [CVisSpace, CImageSpace, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_curv_matrix_csr.argtypes = [CImageSpace, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_grad_matrix_csr.argtypes = [CImageSpace, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_csr_matvec.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int]
        L.oracle_num_threads.restype = C.c_int
    return _lib


def ref_lib():
    """The reference's own deflect.c + raster.c, if oracle/Makefile built them."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, '_ref', 'libref_lens.so')
        if not os.path.exists(path):
            return None
        _ref = _load(path)
        _ref.deflect_points.argtypes = [CLensModel, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        _ref.deflect_points.restype = None
        _ref.manifold_lens_matrix_csr.argtypes = [CImageSpace, CImageSpace, CLensModel, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]
        _ref.manifold_lens_matrix_csr.restype = None
        _ref.triangle_gradient_csr.argtypes = [CDelaunaySpace, C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.triangle_gradient_csr.restype = None
        _ref.delaunay_lens_matrix_csr.argtypes = [CImageSpace, CDelaunaySpace, C.c_int, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.delaunay_lens_matrix_csr.restype = None
        _ref.triangle_geometry_from_inds.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _ref.triangle_geometry_from_inds.restype = None
    return _ref


# ---------------------------------------------------------------------------
# lens parameter packing (lentils/models/lens_models.py:102-116, 129-148)
# ---------------------------------------------------------------------------

def pack_pemd(b, q, th, qh, x, y, rc=0.0, z=0.5):
    r = np.zeros(1, dtype=MASS_MODEL_DTYPE)[0]
    r['type'] = 0
    r['z_l'] = z
    r['fpars'][:9] = [x, y, q, th, np.sin((th + 90) * np.pi / 180), np.cos((th + 90) * np.pi / 180), b, qh, rc]
    return r


def pack_shear(ss, sa, x=0.0, y=0.0, z=0.5, gks=0.0, gka=0.0, gss=0.0, gsa=0.0):
    r = np.zeros(1, dtype=MASS_MODEL_DTYPE)[0]
    r['type'] = 1
    r['z_l'] = z
    f = r['fpars']
    f[0], f[1], f[2], f[3] = x, y, ss, sa
    f[4], f[5] = np.sin((sa + 90) * np.pi / 180), np.cos((sa + 90) * np.pi / 180)
    f[6], f[7] = gks, gka
    f[8], f[9] = np.sin((gka + 90) * np.pi / 180), np.cos((gka + 90) * np.pi / 180)
    f[10], f[11] = gss, gsa
    f[12], f[13] = np.sin((gsa + 90) * np.pi / 180), np.cos((gsa + 90) * np.pi / 180)
    return r


def _lens_struct(lenses):
    arr = np.ascontiguousarray(np.array(lenses, dtype=MASS_MODEL_DTYPE).reshape(-1))
    return CLensModel(len(arr), 1, _ptr(arr)), arr


# ---------------------------------------------------------------------------
# deflection / raster
# ---------------------------------------------------------------------------

def deflect_points(lenses, pts, use_ref=False):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.zeros_like(pts)
    glm, keep = _lens_struct(lenses)
    n = pts.size // 2
    if use_ref:
        R = ref_lib()
        R.deflect_points(glm, _ptr(pts), n, _ptr(out), 0, None)
    else:
        if lib().oracle_deflect_points(glm, _ptr(pts), n, _ptr(out)):
            raise ValueError('unsupported lens type')
    return out


def deflect_points_deriv(lenses, pts, use_ref=False):
    """deflect_points with want_deriv = 1 (deflect.c:60-66, 304-338):
    (deflected points, deriv[12], per_point) where deriv holds the
    derivatives d alpha / d(b, qh, f, x, y, th) as (x, y) pairs summed over
    every point and PEMD component in order (the reference adds them all
    into one array), and per_point [n][12] each point's own sums (None for
    the reference build)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.zeros_like(pts)
    glm, keep = _lens_struct(lenses)
    n = pts.size // 2
    deriv = np.zeros(16)
    if use_ref:
        ref_lib().deflect_points(glm, _ptr(pts), n, _ptr(out), 1, _ptr(deriv))
        return out, deriv[:12].copy(), None
    per = np.zeros((n, 12))
    if lib().oracle_deflect_points_deriv(glm, _ptr(pts), n, _ptr(out), _ptr(deriv), _ptr(per)):
        raise ValueError('unsupported lens type')
    return out, deriv[:12].copy(), per


def stencil_lattice(im):
    pts = np.zeros((2 * im.nx + 1, 2 * im.ny + 1, 2))
    lib().oracle_stencil_lattice(im.cstruct(), _ptr(pts))
    return pts


def lens_matrix_csr(im, src, lenses):
    """Returns (row_inds, cols, vals, status) of the lens matrix."""
    glm, keep = _lens_struct(lenses)
    nnz = C.c_long(0)
    st = lib().oracle_lens_matrix_csr(im.cstruct(), src.cstruct(), glm, None, None, None, C.byref(nnz))
    rp = np.zeros(im.nx * im.ny + 1, dtype=np.int32)
    ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
    va = np.zeros(max(nnz.value, 1), dtype=np.float64)
    lib().oracle_lens_matrix_csr(im.cstruct(), src.cstruct(), glm, _ptr(rp), _ptr(ci), _ptr(va), C.byref(nnz))
    return rp, ci[:nnz.value], va[:nnz.value], st


def lens_matrix_from_lattice(im, src, lattice):
    lattice = np.ascontiguousarray(lattice, dtype=np.float64)
    nnz = C.c_long(0)
    lib().oracle_lens_matrix_from_lattice(im.cstruct(), src.cstruct(), _ptr(lattice), None, None, None,
                                          C.byref(nnz))
    rp = np.zeros(im.nx * im.ny + 1, dtype=np.int32)
    ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
    va = np.zeros(max(nnz.value, 1), dtype=np.float64)
    st = lib().oracle_lens_matrix_from_lattice(im.cstruct(), src.cstruct(), _ptr(lattice), _ptr(rp), _ptr(ci),
                                               _ptr(va), C.byref(nnz))
    return rp, ci[:nnz.value], va[:nnz.value], st


def lens_matrix_from_lattice_nodrop(im, src, lattice):
    lattice = np.ascontiguousarray(lattice, dtype=np.float64)
    nnz = C.c_long(0)
    f = lib().oracle_lens_matrix_from_lattice_nodrop
    f(im.cstruct(), src.cstruct(), _ptr(lattice), None, None, None, C.byref(nnz))
    rp = np.zeros(im.nx * im.ny + 1, dtype=np.int32)
    ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
    va = np.zeros(max(nnz.value, 1), dtype=np.float64)
    st = f(im.cstruct(), src.cstruct(), _ptr(lattice), _ptr(rp), _ptr(ci), _ptr(va), C.byref(nnz))
    return rp, ci[:nnz.value], va[:nnz.value], st


def ref_lens_matrix_csr(im, src, lenses, capacity_per_row=50):
    """The reference's manifold_lens_matrix_csr (raster.c:272), with the
    caller-owned buffers sized like lens_operators.py:30-34."""
    R = ref_lib()
    glm, keep = _lens_struct(lenses)
    n = im.nx * im.ny
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(capacity_per_row * n, dtype=np.int32)
    va = np.zeros(capacity_per_row * n, dtype=np.float64)
    R.manifold_lens_matrix_csr(im.cstruct(), src.cstruct(), glm, _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci[:rp[-1]], va[:rp[-1]]


# ---------------------------------------------------------------------------
# NUFFT (lentils/operators/fourier_operators.py, lentils/common/space_base.py)
# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# Delaunay source planes (lentils/backend/triangles.c)
# ---------------------------------------------------------------------------

def _dspace(pts, tris):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    return CDelaunaySpace(pts.shape[0], tris.shape[0], 1, 1, None, _ptr(pts), _ptr(tris)), (pts, tris)


def triangle_gradient_csr(pts, tris, use_ref=False):
    """(row_inds, cols, vals) of the weighted gradient rows, two per triangle."""
    ds, (p, t) = _dspace(pts, tris)
    nt = t.shape[0]
    rp = np.zeros(2 * nt + 1, dtype=np.int32)
    ci = np.zeros(6 * nt, dtype=np.int32)
    va = np.zeros(6 * nt)
    if use_ref:
        ref_lib().triangle_gradient_csr(ds, _ptr(rp), _ptr(ci), _ptr(va))
    else:
        lib().oracle_triangle_gradient_csr(nt, _ptr(p), _ptr(t), _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci, va


def delaunay_lens_matrix_csr(im, pts, tris, ncast, unc_pts, unc_tri, use_ref=False):
    """(row_inds, cols, vals) of the Delaunay lens matrix (triangles.c:106-145);
    unc_pts / unc_tri: the uncasted pixels' deflected positions and their
    triangle indices (-1 outside), in x-major pixel order."""
    ds, (p, t) = _dspace(pts, tris)
    up = np.ascontiguousarray(unc_pts, dtype=np.float64).reshape(-1, 2)
    ut = np.ascontiguousarray(unc_tri, dtype=np.int32).reshape(-1)
    n = im.nx * im.ny
    cap = p.shape[0] + 3 * up.shape[0]
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(max(cap, 1), dtype=np.int32)
    va = np.zeros(max(cap, 1))
    if use_ref:
        ref_lib().delaunay_lens_matrix_csr(im.cstruct(), ds, int(ncast), _ptr(up), _ptr(ut), _ptr(rp), _ptr(ci),
                                           _ptr(va))
    else:
        lib().oracle_delaunay_lens_matrix_csr(im.cstruct(), t.shape[0], _ptr(p), _ptr(t), int(ncast), _ptr(up),
                                              _ptr(ut), _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci[:rp[-1]], va[:rp[-1]]


def bessel_i0(x):
    return lib().oracle_bessel_i0(float(x))


def kb_beta_default(wsup, pad):
    """fourier_operators.py:108"""
    return np.pi * np.sqrt((2.0 * wsup / pad) ** 2 * (pad - 0.5) ** 2 - 0.8)


class NUFFTGeometry:
    """Padded space, Fourier grid and apodization of NUFFTOperator
    (fourier_operators.py:27-68, 96-126; space_base.py:62-84)."""

    def __init__(self, im, pad=2, wsup=4, beta=None):
        self.im = im
        self.pad = int(pad)
        self.wsup = int(wsup)
        self.beta = kb_beta_default(wsup, pad) if beta is None else float(beta)
        self.padx = ((self.pad - 1) * im.nx + 1) // 2
        self.pady = ((self.pad - 1) * im.ny + 1) // 2
        self.nxp, self.nyp = im.nx + 2 * self.padx, im.ny + 2 * self.pady
        b = im.bounds + np.array([-self.padx * im.dx, self.padx * im.dx, -self.pady * im.dy, self.pady * im.dy])
        self.pbounds = b
        self.nu, self.nv = self.nxp, self.nyp
        self.half_nv = self.nv // 2 + 1
        lx = ARCSEC * (b[1] - b[0])
        ly = ARCSEC * (b[3] - b[2])
        self.du, self.dv = 1.0 / lx, 1.0 / ly
        self.gcx = ARCSEC * 0.5 * (b[1] + b[0])
        self.gcy = ARCSEC * 0.5 * (b[3] + b[2])
        cx = (np.arange(im.nx) - 0.5 * im.nx) / self.nxp
        cy = (np.arange(im.ny) - 0.5 * im.ny) / self.nyp
        ax = np.sqrt(self.beta ** 2 - (2 * np.pi * self.wsup * cx) ** 2)
        ay = np.sqrt(self.beta ** 2 - (2 * np.pi * self.wsup * cy) ** 2)
        self.apod = 1.0 / np.outer(np.sinh(ax) / ax, np.sinh(ay) / ay)
        self._ch = np.zeros(1)

    def fstruct(self):
        return CFourierSpace(self.nu, self.nv, self.half_nv, 1, 1, self.du, self.dv, self.gcx, self.gcy,
                             _ptr(self._ch))


def vis_struct(uvw, channels):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64).reshape(-1, 3)
    ch = np.ascontiguousarray(channels, dtype=np.float64).reshape(-1)
    return CVisSpace(uvw.shape[0], ch.size, 1, _ptr(uvw), _ptr(ch)), (uvw, ch)


def grid(uvw, channels, vis, geo, direction, grid_in=None, nthreads=1):
    """grid_cpu restated: direction 0 degrid (returns vis), 1 grid (returns grid)."""
    vs, keep = vis_struct(uvw, channels)
    fs = geo.fstruct()
    if direction == 0:
        g = np.ascontiguousarray(grid_in, dtype=np.complex128)
        out = np.zeros((keep[1].size, keep[0].shape[0]), dtype=np.complex128)
        lib().oracle_grid(vs, _ptr(out), fs, _ptr(g), geo.wsup, geo.beta, 0, nthreads)
        return out
    v = np.ascontiguousarray(vis, dtype=np.complex128)
    out = np.zeros((geo.nu, geo.half_nv), dtype=np.complex128)
    lib().oracle_grid(vs, _ptr(v), fs, _ptr(out), geo.wsup, geo.beta, 1, nthreads)
    return out


def nufft_forward(img, uvw, channels, geo, nthreads=1):
    """gridder * fft * zpad * apod applied to a real image (nx, ny)."""
    x = np.asarray(img, dtype=np.float64).reshape(geo.im.nx, geo.im.ny) * geo.apod
    p = np.zeros((geo.nxp, geo.nyp))
    p[geo.padx:geo.padx + geo.im.nx, geo.pady:geo.pady + geo.im.ny] = x
    g = np.fft.rfft2(p, norm='backward')
    return grid(uvw, channels, None, geo, 0, grid_in=g, nthreads=nthreads)


def nufft_adjoint(vis, uvw, channels, geo, nthreads=1):
    g = grid(uvw, channels, vis, geo, 1, nthreads=nthreads)
    p = np.fft.irfft2(g, s=(geo.nxp, geo.nyp), norm='forward')
    x = p[geo.padx:geo.padx + geo.im.nx, geo.pady:geo.pady + geo.im.ny]
    return x * geo.apod


def beam_grid(im):
    """The dirty beam's image space of RadioDataset._make_beam_space
    (data_base.py:148-160): twice the image in each direction, bounds
    +-dx*n with the x pixel size for both axes."""
    rx = im.dx * im.nx
    ry = im.dx * im.ny
    return Grid2D(2 * im.nx, 2 * im.ny, (-rx, rx, -ry, ry))


def dirty_beam(uvw, channels, im, weights, wsup=4, beta=None, nthreads=1):
    """RadioDataset.dirty_beam (data_base.py:163-172): the NUFFT adjoint
    (pad 2; the reference's defaults W = 4 and Beatty's beta) of C 1 on the
    beam space."""
    bs = beam_grid(im)
    geo = NUFFTGeometry(bs, pad=2, wsup=wsup, beta=beta)
    c1 = np.asarray(weights, dtype=np.float64).reshape(1, -1).astype(np.complex128)
    return nufft_adjoint(c1, uvw, channels, geo, nthreads=nthreads)


def bcb_apply(img, beam, geo):
    """RadioDataset.blurred_covariance_operator (data_base.py:115-129):
    zpad^T fft^T diag(rfft2(roll(beam, (nxp/2, nyp/2))) / N) fft zpad, with
    geo the main NUFFT's geometry (its padded grid must be the beam's)."""
    db = np.roll(beam, [geo.nxp // 2, geo.nyp // 2], axis=[0, 1])
    dbf = np.fft.rfft2(db) / db.size
    p = np.zeros((geo.nxp, geo.nyp))
    p[geo.padx:geo.padx + geo.im.nx, geo.pady:geo.pady + geo.im.ny] = np.asarray(img).reshape(geo.im.nx, geo.im.ny)
    q = np.fft.irfft2(dbf * np.fft.rfft2(p), s=p.shape, norm='forward')
    return q[geo.padx:geo.padx + geo.im.nx, geo.pady:geo.pady + geo.im.ny]


def dft_matrix(uvw, channels, im):
    vs, keep = vis_struct(uvw, channels)
    nm = int(im.mask.sum())
    n = keep[0].shape[0]
    rp = np.zeros(2 * n + 1, dtype=np.int32)
    ci = np.zeros(2 * n * nm, dtype=np.int32)
    va = np.zeros(2 * n * nm, dtype=np.float64)
    lib().oracle_dft_matrix_csr(vs, im.cstruct(), _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci, va


def curv_matrix(im):
    n = im.nx * im.ny
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(5 * n, dtype=np.int32)
    va = np.zeros(5 * n, dtype=np.float64)
    lib().oracle_curv_matrix_csr(im.cstruct(), _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci[:rp[-1]], va[:rp[-1]]


def grad_matrix(im):
    n = im.nx * im.ny
    rp = np.zeros(2 * n + 1, dtype=np.int32)
    ci = np.zeros(4 * n, dtype=np.int32)
    va = np.zeros(4 * n, dtype=np.float64)
    lib().oracle_grad_matrix_csr(im.cstruct(), _ptr(rp), _ptr(ci), _ptr(va))
    return rp, ci[:rp[-1]], va[:rp[-1]]


def csr_matvec(rp, ci, va, x, nthreads=1):
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    va = np.ascontiguousarray(va, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(rp.size - 1)
    lib().oracle_csr_matvec(rp.size - 1, _ptr(rp), _ptr(ci), _ptr(va), _ptr(x), _ptr(y), nthreads)
    return y


def max_relative_error(a, b):
    """max|a-b| / max|b|, lentils/tests/test_utils.py:13-20 ('max')."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
