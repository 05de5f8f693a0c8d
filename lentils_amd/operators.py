"""Operator algebra and the lensing / Fourier operators of Lentils.

Drop-in counterparts of lentils/operators/{operator_base,lens_operators,
fourier_operators}.py: same class names, constructor arguments, space
shapes, ``apply`` / ``*`` / ``+`` / ``.T`` semantics and error behaviour
(ValueError on shape mismatch and on the Nyquist guard).  The arithmetic runs
in liblentils_hip.so on the GPU:

* explicit-matrix operators apply through the library's CSR SpMV,
* ManifoldLensOperator builds its CSR matrix on the GPU
  (deflection lattice + exact rasterization),
* NUFFTOperator / GriddingOperator / FFTOperator call the HIP gridder,
  degridder and rocFFT.

Zero padding and diagonal scalings are data movement / elementwise scaling
on the host arrays the caller passes in.  For device-resident inversions use
lentils_amd.solvers.
"""
import ctypes as C
from copy import copy

import numpy as np
from scipy import sparse

from . import backend as B
from .spaces import DelaunaySpace, FourierSpace, ImageSpace, VisibilitySpace


def broadcastable(a, b):
    return all((m == n) or (m == 1) or (n == 1) for m, n in zip(a.shape[::-1], b.shape[::-1]))


def _gpu_csr_matvec(mat, x):
    """y = mat @ x on the GPU for a real scipy sparse matrix."""
    csr = mat if sparse.isspmatrix_csr(mat) else mat.tocsr()
    rp = np.ascontiguousarray(csr.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(csr.indices, dtype=np.int32)
    va = np.ascontiguousarray(csr.data, dtype=np.float64)
    xv = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    if xv.size != csr.shape[1]:
        raise ValueError(f'Incompatible shapes! matrix {csr.shape}, vector {xv.size}')
    y = np.zeros(csr.shape[0])
    B.check(B.lib().lt_csr_spmv_host(csr.shape[0], csr.shape[1], B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(xv),
                                     B.ptr(y)))
    return y


class Operator:
    """Linear map space_right -> space_left (operator_base.py:46-113)."""

    def __init__(self, space_left, space_right):
        self.space_right = space_right
        self.space_left = space_left

    def _make_transpose(self):
        tmp = copy(self)
        tmp.space_left, tmp.space_right = tmp.space_right, tmp.space_left
        try:
            tmp._mat = tmp._mat.T
            tmp._csr_cache = None
        except AttributeError:
            tmp._matrixfree_forward, tmp._matrixfree_transpose = tmp._matrixfree_transpose, tmp._matrixfree_forward
        return tmp

    @property
    def matrix(self):
        return self._mat if self.has_matrix else None

    @property
    def has_matrix(self):
        return hasattr(self, '_mat')

    @property
    def T(self):
        try:
            return self._transpose
        except AttributeError:
            self._transpose = self._make_transpose()
            self._transpose._transpose = self
            return self._transpose

    def __mul__(self, other):
        return self.apply(other)

    def __add__(self, other):
        return CompositeOperatorSum([self, other])

    def _cast_output(self, vec):
        return np.ascontiguousarray(vec).view(self.space_left.dtype).reshape(self.space_left.shape)

    def _matrix_apply(self, other):
        mat = self._mat
        v = other.view(mat.dtype).flatten() if other.dtype != mat.dtype else other.flatten()
        if np.iscomplexobj(mat.data) or np.iscomplexobj(v):
            # complex explicit matrices only occur as diagonals here
            return (mat @ v)
        cache = getattr(self, '_csr_cache', None)
        if cache is None:
            cache = mat.tocsr()
            self._csr_cache = cache
        return _gpu_csr_matvec(cache, v)

    def apply(self, other):
        if isinstance(other, np.ndarray):
            if not broadcastable(self.space_right, other.reshape((-1,) + self.space_right.shape[-2:])):
                raise ValueError(f'Incompatible shapes! {self.space_right.shape}, {other.shape}')
            if self.has_matrix:
                out = self._matrix_apply(other)
            else:
                out = self._matrixfree_forward(other)
            return self._cast_output(out)
        elif isinstance(other, Operator):
            return CompositeOperatorProduct([self, other])
        raise TypeError('Operator.apply() can only be used on numpy.ndarray or Operator')

    def _matrixfree_forward(self, other):
        raise NotImplementedError

    def _matrixfree_transpose(self, other):
        raise NotImplementedError

    def __repr__(self):
        return f'<{type(self).__name__}: {self.space_right} -> {self.space_left}>'


class CompositeOperatorProduct(Operator):
    """Product of operators applied right to left (operator_base.py:116-143)."""

    def __init__(self, oplist):
        self.subops = []
        for op in oplist:
            if isinstance(op, CompositeOperatorProduct):
                self.subops += op.subops
            else:
                self.subops.append(op)
        for op1, op2 in zip(self.subops[:-1], self.subops[1:]):
            if not broadcastable(op1.space_right, op2.space_left):
                raise ValueError(f'Incompatible shapes! {op1.space_right.shape}, {op2.space_left.shape}')
        super().__init__(oplist[0].space_left, oplist[-1].space_right)

    def _make_transpose(self):
        return CompositeOperatorProduct([op.T for op in self.subops[::-1]])

    def _matrixfree_forward(self, vec):
        out = vec
        for op in self.subops[::-1]:
            out = op * out
        return out


class CompositeOperatorSum(Operator):
    """Sum of operators on the same spaces (operator_base.py:146-175)."""

    def __init__(self, oplist):
        self.subops = []
        for op in oplist:
            if isinstance(op, CompositeOperatorSum):
                self.subops += op.subops
            else:
                self.subops.append(op)
        for op1, op2 in zip(self.subops[:-1], self.subops[1:]):
            if not broadcastable(op1.space_right, op2.space_right):
                raise ValueError(f'Incompatible shapes! {op1.space_right.shape}, {op2.space_right.shape}')
            if not broadcastable(op1.space_left, op2.space_left):
                raise ValueError(f'Incompatible shapes! {op1.space_left.shape}, {op2.space_left.shape}')
        super().__init__(oplist[0].space_left, oplist[-1].space_right)

    def _make_transpose(self):
        return CompositeOperatorSum([op.T for op in self.subops])

    def _matrixfree_forward(self, vec):
        out = self.subops[0] * vec
        for op in self.subops[1:]:
            out = out + op * vec
        return out


class DiagonalOperator(Operator):
    """diag(data) on one space; options='r2c' duplicates real entries over
    (re, im) pairs (operator_base.py:178-197)."""

    def __init__(self, space, data, options=None, **superargs):
        if options is None:
            self._data = np.asarray(data)
            matdim = int(np.prod(space.shape))
        elif options == 'r2c':
            flatreal = np.asarray(data).flatten().real
            matdim = 2 * flatreal.size
            self._data = np.zeros(matdim, dtype=np.float64)
            self._data[0::2] = flatreal
            self._data[1::2] = flatreal
        else:
            raise ValueError('options should be either None or r2c')
        self._mat = sparse.spdiags(self._data.flatten(), 0, matdim, matdim)
        super().__init__(space, space)

    def _matrix_apply(self, other):
        d = self._data.flatten()
        return d * other.view(d.dtype).flatten()


class PriorCovarianceOperator(Operator):
    """lambda * R^T R for R = identity ('tikhonov'), forward differences
    ('gradient') or the 5-point Laplacian ('curvature'); the finite-difference
    matrices are built on the GPU (operator_base.py:200-251)."""

    def __init__(self, space, type, strength, lambda_boundary=None, **superargs):
        n = space.size
        if isinstance(space, DelaunaySpace):
            # triangle meshes: identity or the weighted triangle gradients
            # (triangle_gradient_csr, triangles.c:66-103, on the GPU)
            if type == 'tikhonov':
                self._op = sparse.spdiags(np.ones(n), 0, n, n)
            elif type == 'gradient':
                nt = space.num_triangles
                rp = np.zeros(2 * nt + 1, dtype=np.int32)
                ci = np.zeros(max(6 * nt, 1), dtype=np.int32)
                va = np.zeros(max(6 * nt, 1), dtype=np.float64)
                B.check(B.lib().lt_triangle_gradient_csr(space, B.ptr(rp), B.ptr(ci), B.ptr(va)))
                self._op = sparse.csr_matrix((va[:6 * nt], ci[:6 * nt], rp), shape=(2 * nt, n))
            else:
                raise NotImplementedError('Curvature regularization on triangles not yet implemented')
            self.type = type
            self._lambda = strength
            self._mat = (self._lambda * self._op.T.dot(self._op)).tocsr()
            if lambda_boundary is not None:
                data = np.zeros(n)
                data[space.edge] = strength * lambda_boundary
                self._mat = (self._mat + sparse.spdiags(data, 0, n, n)).tocsr()
            self._transpose = self
            Operator.__init__(self, space, space)
            return
        if not isinstance(space, ImageSpace):
            raise NotImplementedError('PriorCovarianceOperator not implement for this space')
        if type == 'tikhonov':
            self._op = sparse.spdiags(np.ones(n), 0, n, n)
        elif type == 'gradient':
            rp = np.zeros(2 * n + 1, dtype=np.int32)
            ci = np.zeros(4 * n, dtype=np.int32)
            va = np.zeros(4 * n, dtype=np.float64)
            B.lib().grad_matrix_findiff_csr(space, B.ptr(rp), B.ptr(ci), B.ptr(va))
            self._op = sparse.csr_matrix((va[:rp[-1]], ci[:rp[-1]], rp), shape=(2 * n, n))
        elif type == 'curvature':
            rp = np.zeros(n + 1, dtype=np.int32)
            ci = np.zeros(5 * n, dtype=np.int32)
            va = np.zeros(5 * n, dtype=np.float64)
            B.lib().curv_matrix_findiff_csr(space, B.ptr(rp), B.ptr(ci), B.ptr(va))
            self._op = sparse.csr_matrix((va[:rp[-1]], ci[:rp[-1]], rp), shape=(n, n))
        else:
            raise ValueError(f'unknown prior type {type!r}')
        self.type = type
        self._lambda = strength
        self._mat = (self._lambda * self._op.T.dot(self._op)).tocsr()
        self._transpose = self
        super().__init__(space, space)


class LensOperator(Operator):

    def __init__(self, space_left, space_right, lensmodel):
        self._lensmodel = lensmodel
        super().__init__(space_left, space_right)


class ManifoldLensOperator(LensOperator):
    """CSR lens matrix image -> source built on the GPU: FP64 deflection of
    the stencil lattice, exact triangle rasterization, and the reference's
    reduction (lens_operators.py:22-40, raster.c:272-395).  The exact entry
    count replaces the reference's 50-per-row capacity guess.
    ``raster_status`` carries the reference's overflow warnings as bits."""

    def __init__(self, image_space, source_space, lensmodel, z_src, **superargs):
        lensmodel.setup_raytracing(z_src)
        n = image_space.size
        nnz = C.c_long(0)
        st = C.c_int(0)
        L = B.lib()
        B.check(L.lt_lens_matrix_nnz(image_space, source_space, lensmodel, C.byref(nnz), C.byref(st)))
        rp = np.zeros(n + 1, dtype=np.int32)
        ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        va = np.zeros(max(nnz.value, 1), dtype=np.float64)
        B.check(L.lt_lens_matrix_csr(image_space, source_space, lensmodel, B.ptr(rp), B.ptr(ci), B.ptr(va),
                                     ci.size, C.byref(nnz), C.byref(st)))
        self.raster_status = st.value
        self._mat = sparse.csr_matrix((va[:nnz.value], ci[:nnz.value], rp), shape=(n, source_space.size))
        super().__init__(image_space, source_space, lensmodel)


class DelaunayLensOperator(LensOperator):
    """Lens matrix onto an adaptive, triangulated source plane (Vegetti &
    Koopmans; lens_operators.py:45-84): every ncasted-th image pixel (in
    both directions, inside the mask) is ray-traced to a vertex of the
    source triangulation, every other unmasked pixel is interpolated
    linearly in the triangle its deflected position falls in.  Deflection
    and the CSR assembly run on the GPU; the triangulation and the point
    location are scipy's, as in the reference."""

    def __init__(self, image_space, lensmodel, z_src, ncasted=1, **superargs):
        from .models import GlobalLensModel
        if not isinstance(image_space, ImageSpace):
            raise TypeError('image_space must be ImageSpace')
        if not isinstance(lensmodel, GlobalLensModel):
            raise TypeError('lensmodel must be GlobalLensModel')
        ncasted = int(ncasted)
        image_mask = np.asarray(image_space.mask, dtype=bool)
        image_points = image_space.points
        casted_mask = np.zeros_like(image_mask)
        casted_mask[::ncasted, ::ncasted] = True
        uncasted_mask = ~casted_mask
        casted_mask &= image_mask
        uncasted_mask &= image_mask
        self.casted_points = lensmodel.deflect(np.ascontiguousarray(image_points[casted_mask]), z_s=z_src)
        self.uncasted_points = np.ascontiguousarray(
            lensmodel.deflect(np.ascontiguousarray(image_points[uncasted_mask]), z_s=z_src))
        source_space = DelaunaySpace(self.casted_points)
        uncasted_tri_inds = np.ascontiguousarray(source_space.delaunay.find_simplex(self.uncasted_points),
                                                 dtype=np.int32)
        num_rows = image_space.size
        num_vals = int(np.sum(casted_mask)) + 3 * int(np.sum(uncasted_mask))
        rp = np.zeros(num_rows + 1, dtype=np.int32)
        ci = np.zeros(max(num_vals, 1), dtype=np.int32)
        va = np.zeros(max(num_vals, 1), dtype=np.float64)
        nnz = C.c_long(0)
        B.check(B.lib().lt_delaunay_lens_matrix_csr(image_space, source_space, ncasted, B.ptr(self.uncasted_points),
                                                    B.ptr(uncasted_tri_inds), B.ptr(rp), B.ptr(ci), B.ptr(va),
                                                    ci.size, C.byref(nnz)))
        self._mat = sparse.csr_matrix((va[:nnz.value], ci[:nnz.value], rp), shape=(num_rows, source_space.size))
        super().__init__(image_space, source_space, lensmodel)


class FFTOperator(Operator):
    """rfft2 (unnormalised) image -> FourierSpace; transpose irfft2 with
    norm='forward' (fourier_operators.py:12-24), on rocFFT."""

    def __init__(self, image_space, **superargs):
        super().__init__(FourierSpace(image_space), image_space)

    def _fft(self, vec, forward):
        img_space = self.space_right if forward else self.space_left
        nx, ny = img_space.shape[-2:]
        if forward:
            x = np.ascontiguousarray(vec, dtype=np.float64).reshape(-1, nx, ny)
            out = np.zeros((x.shape[0], nx, ny // 2 + 1), dtype=np.complex128)
            B.check(B.lib().lt_fft_r2c_host(nx, ny, x.shape[0], B.ptr(x), B.ptr(out)))
        else:
            x = np.ascontiguousarray(vec, dtype=np.complex128).reshape(-1, nx, ny // 2 + 1)
            out = np.zeros((x.shape[0], nx, ny))
            B.check(B.lib().lt_fft_c2r_host(nx, ny, x.shape[0], B.ptr(x), B.ptr(out)))
        return out

    def _matrixfree_forward(self, vec):
        return self._cast_output(self._fft(vec, True))

    def _matrixfree_transpose(self, vec):
        return self._cast_output(self._fft(vec, False))


class ZeroPaddingOperator(Operator):
    """Symmetric zero padding by ((pad-1) n + 1) // 2 per side; transpose
    crops (fourier_operators.py:27-52)."""

    def __init__(self, image_space, pad_factor, **superargs):
        self.pad_factor = int(pad_factor)
        self.padx = ((self.pad_factor - 1) * image_space.nx + 1) // 2
        self.pady = ((self.pad_factor - 1) * image_space.ny + 1) // 2
        shape = (image_space.nx + 2 * self.padx, image_space.ny + 2 * self.pady)
        dx, dy = image_space.dx, image_space.dy
        bounds = image_space.bounds.copy()
        bounds += np.array([-self.padx * dx, self.padx * dx, -self.pady * dy, self.pady * dy])
        padded_space = ImageSpace(shape=shape, bounds=bounds, channels=image_space.channels, mask=None)
        super().__init__(padded_space, image_space)

    def _matrixfree_forward(self, vec):
        out = self.space_left.new_vector()
        out[..., self.padx:-self.padx, self.pady:-self.pady] = vec[..., :, :]
        return out

    def _matrixfree_transpose(self, vec):
        out = self.space_left.new_vector()
        out[..., :, :] = vec[..., self.padx:-self.padx, self.pady:-self.pady]
        return out


class ApodizationCorrectionOperator(DiagonalOperator):
    """1 / (sinhc x sinhc) grid-correction of the KB kernel (Beatty et al.
    2005; fourier_operators.py:54-68)."""

    def __init__(self, image_space, padded_space, kb_beta, kernel_support, **superargs):
        self.kernel_support = int(kernel_support)
        self.kb_beta = float(kb_beta)
        cx = (np.arange(image_space.nx) - 0.5 * image_space.nx) / padded_space.nx
        cy = (np.arange(image_space.ny) - 0.5 * image_space.ny) / padded_space.ny
        argx = np.sqrt(self.kb_beta ** 2 - (2 * np.pi * self.kernel_support * cx) ** 2)
        argy = np.sqrt(self.kb_beta ** 2 - (2 * np.pi * self.kernel_support * cy) ** 2)
        apod_data = 1.0 / np.outer(np.sinh(argx) / argx, np.sinh(argy) / argy)
        super().__init__(image_space, apod_data)


def kb_beta_default(kernel_support, pad_factor):
    """Beatty et al. (2005) optimal beta, fourier_operators.py:108."""
    return np.pi * np.sqrt((2.0 * kernel_support / pad_factor) ** 2 * (pad_factor - 0.5) ** 2 - 0.8)


class GriddingOperator(Operator):
    """KB gridding on the GPU: forward degrids FourierSpace -> visibilities,
    transpose grids (fourier_operators.py:71-93)."""

    def __init__(self, space_vis, space_fourier, kb_beta, kernel_support, **superargs):
        self.kernel_support = int(kernel_support)
        self.kb_beta = float(kb_beta)
        iumax = np.max(space_vis.channels) * np.max(np.abs(space_vis.uvw[:, 0])) / space_fourier.du
        ivmax = np.max(space_vis.channels) * np.max(np.abs(space_vis.uvw[:, 1])) / space_fourier.dv
        if iumax + self.kernel_support + 1 > 0.5 * space_fourier.nu or \
                ivmax + self.kernel_support + 1 > 0.5 * space_fourier.nv:
            raise ValueError('Visibility space contains points higher than the Nyquist frequency of the grid.')
        super().__init__(space_vis, space_fourier)

    def _matrixfree_forward(self, vec):
        out = self.space_left.new_vector()
        g = np.ascontiguousarray(vec, dtype=np.complex128)
        B.lib().grid_cpu(self.space_left, B.ptr(out), self.space_right, B.ptr(g), self.kernel_support,
                         self.kb_beta, 0)
        return out

    def _matrixfree_transpose(self, vec):
        out = self.space_left.new_vector()
        v = np.ascontiguousarray(vec, dtype=np.complex128)
        B.lib().grid_cpu(self.space_right, B.ptr(v), self.space_left, B.ptr(out), self.kernel_support,
                         self.kb_beta, 1)
        return out


class NUFFTOperator(Operator):
    """image -> visibilities: gridder * fft * zpad * apod
    (fourier_operators.py:96-126), applied in one GPU pass per call through a
    device-resident plan (uv set binned into uv tiles once).

    Extra keyword arguments beyond the reference: ``kb_beta`` (explicit KB
    beta instead of Beatty's formula) and ``precision`` ('f64' or 'f32' for
    the gridding arithmetic)."""

    def __init__(self, vis_space, image_space, pad_factor=2, kernel_support=4, combine_channels=True,
                 combine_stokes=True, kb_beta=None, precision='f64', **superargs):
        self.image_space = image_space
        self.vis_space = vis_space
        self.pad_factor = int(pad_factor)
        self.kernel_support = int(kernel_support)
        self.kb_beta = kb_beta_default(self.kernel_support, self.pad_factor) if kb_beta is None else float(kb_beta)
        self.precision = precision
        self.zpad = ZeroPaddingOperator(self.image_space, self.pad_factor)
        self.padded_space = self.zpad.space_left
        self.apod = ApodizationCorrectionOperator(self.image_space, self.padded_space, self.kb_beta,
                                                  self.kernel_support)
        self.fft = FFTOperator(self.padded_space)
        self.fourier_space = self.fft.space_left
        self.gridder = GriddingOperator(vis_space, self.fourier_space, self.kb_beta, self.kernel_support)
        self.subops = [self.gridder, self.fft, self.zpad, self.apod]
        prec = B.LT_F64 if precision == 'f64' else B.LT_F32
        h = C.c_void_p()
        B.check(B.lib().lt_nufft_create(vis_space, image_space, self.pad_factor, self.kernel_support,
                                        self.kb_beta, prec, C.byref(h)))
        self._handle = h
        self._is_adjoint = False
        super().__init__(vis_space, image_space)

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and getattr(self, '_owner', True) and B._lib is not None:
            B._lib.lt_nufft_destroy(h)
            self._handle = None

    def _make_transpose(self):
        tmp = copy(self)
        tmp._owner = False
        tmp.space_left, tmp.space_right = self.space_right, self.space_left
        tmp._is_adjoint = not self._is_adjoint
        return tmp

    def _fwd(self, vec):
        x = np.ascontiguousarray(vec, dtype=np.float64).reshape(-1)
        out = np.zeros(self.vis_space.shape, dtype=np.complex128)
        B.check(B.lib().lt_nufft_forward_host(self._handle, B.ptr(x), B.ptr(out)))
        return out

    def _adj(self, vec):
        v = np.ascontiguousarray(vec, dtype=np.complex128).reshape(-1)
        out = np.zeros(self.image_space.shape)
        B.check(B.lib().lt_nufft_adjoint_host(self._handle, B.ptr(v), B.ptr(out)))
        return out

    def _matrixfree_forward(self, vec):
        return self._adj(vec) if self._is_adjoint else self._fwd(vec)


class DFTOperator(Operator):
    """Dense direct Fourier transform as a CSR matrix (2 real rows per
    visibility), built on the GPU (fourier_operators.py:130-145)."""

    def __init__(self, vis_space, image_space, **superargs):
        num_rows = 2 * vis_space.size
        nnz_per_row = int(np.sum(image_space.mask))
        num_vals = num_rows * nnz_per_row
        rp = np.zeros(num_rows + 1, dtype=np.int32)
        ci = np.zeros(max(num_vals, 1), dtype=np.int32)
        va = np.zeros(max(num_vals, 1), dtype=np.float64)
        B.lib().dft_matrix_csr(vis_space, image_space, B.ptr(rp), B.ptr(ci), B.ptr(va))
        self._mat = sparse.csr_matrix((va[:num_vals], ci[:num_vals], rp), shape=(num_rows, image_space.size))
        super().__init__(vis_space, image_space)
