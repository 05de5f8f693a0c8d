"""Device-resident source inversion for batches of lens samples.

The reference solves one source per lens model with scipy's cg over its
operator algebra (lentils/tests/test_all.py:57-74):

    lhs = L^T (N^T C N  or  BCB) L + R,   rhs = L^T N^T C d.

Here every sample of a batch is solved on the GPU at once: the lens matrices
are built together (LensBatch), the NUFFT plan is shared, and a batched
FP64 CG runs to ||b - A x|| / ||b|| <= rtol with no host round trip per
iteration.  ``operator='nufft'`` applies N^T C N exactly (fused
degrid-weight-grid); ``operator='bcb'`` uses the dirty-beam convolution
(data_base.py:115-129).
"""
import ctypes as C

import numpy as np

from . import backend as B
from .spaces import ImageSpace


def _image_array(spaces):
    arr = (B.image_space_ctype * len(spaces))()
    for k, s in enumerate(spaces):
        arr[k] = s
    return arr


def _weights(weights, sigma, mask):
    if weights is None:
        s = np.asarray(sigma, dtype=np.float64).reshape(-1)
        m = np.ones(s.shape, dtype=bool) if mask is None else np.asarray(mask, dtype=bool).reshape(-1)
        w = np.zeros(s.shape)
        w[m] = s[m] ** -2  # data_base.py:36-44
    else:
        w = np.asarray(weights, dtype=np.float64).reshape(-1)
    return np.ascontiguousarray(w)


class DeviceArray:
    """float64 array resident in device memory (inputs that stay on the GPU
    between solves, e.g. the dirty images of a batch of observations)."""

    def __init__(self, array):
        a = np.ascontiguousarray(array, dtype=np.float64)
        self.shape, self.size = a.shape, a.size
        p = C.c_void_p()
        B.check(B.lib().lt_device_alloc(a.nbytes, C.byref(p)))
        self.ptr = p
        B.check(B.lib().lt_device_upload(p, B.ptr(a), a.nbytes))

    def numpy(self):
        out = np.zeros(self.shape)
        B.check(B.lib().lt_device_download(self.ptr, B.ptr(out), out.nbytes))
        return out

    def __del__(self):
        p = getattr(self, 'ptr', None)
        if p is not None and B._lib is not None:
            B._lib.lt_device_free(p)
            self.ptr = None


class LensBatch:
    """Lens matrices L_s (image -> source_spaces[s]) for many lens models,
    built in one GPU pass and kept on the device with their transposes."""

    def __init__(self, image_space, source_spaces, lensmodels, z_src):
        if isinstance(source_spaces, ImageSpace):
            source_spaces = [source_spaces] * len(lensmodels)
        if len(source_spaces) != len(lensmodels):
            raise ValueError('one source space per lens model is required')
        self.image_space = image_space
        self.source_spaces = list(source_spaces)
        recs, counts = [], []
        for lm in lensmodels:
            lm.setup_raytracing(z_src)
            recs.append(lm.lenses)
            counts.append(lm.num_lenses)
        self._lenses = np.ascontiguousarray(np.concatenate(recs)) if recs else np.zeros(0, dtype=B.generic_mass_model_ctype)
        if self._lenses.size == 0:
            self._lenses = np.zeros(1, dtype=B.generic_mass_model_ctype)
        self._counts = np.ascontiguousarray(counts, dtype=np.int32)
        self._srcs = _image_array(self.source_spaces)
        h = C.c_void_p()
        B.check(B.lib().lt_lensbatch_create(image_space, C.cast(self._srcs, C.c_void_p), len(lensmodels),
                                            B.ptr(self._lenses), B.ptr(self._counts), C.byref(h)))
        self._handle = h
        self.nsamples = len(lensmodels)

    def info(self, s):
        nnz = C.c_long(0)
        st = C.c_int(0)
        B.check(B.lib().lt_lensbatch_info(self._handle, s, C.byref(nnz), C.byref(st)))
        return nnz.value, st.value

    def csr(self, s):
        """(row_inds, cols, vals) of sample s, as manifold_lens_matrix_csr writes them."""
        nnz, _ = self.info(s)
        rp = np.zeros(self.image_space.size + 1, dtype=np.int32)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        va = np.zeros(max(nnz, 1), dtype=np.float64)
        B.check(B.lib().lt_lensbatch_download(self._handle, s, B.ptr(rp), B.ptr(ci), B.ptr(va)))
        return rp, ci[:nnz], va[:nnz]

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and B._lib is not None:
            B._lib.lt_lensbatch_destroy(h)
            self._handle = None


class NUFFTPlan:
    """Device-resident NUFFT plan (uv tiles, apodization, rocFFT)."""

    def __init__(self, vis_space, image_space, pad_factor=2, kernel_support=4, kb_beta=None, precision='f64'):
        from .operators import kb_beta_default
        self.vis_space = vis_space
        self.image_space = image_space
        self.kernel_support = int(kernel_support)
        self.kb_beta = kb_beta_default(kernel_support, pad_factor) if kb_beta is None else float(kb_beta)
        prec = B.LT_F64 if precision == 'f64' else B.LT_F32
        h = C.c_void_p()
        B.check(B.lib().lt_nufft_create(vis_space, image_space, int(pad_factor), self.kernel_support, self.kb_beta,
                                        prec, C.byref(h)))
        self._handle = h

    def forward(self, image):
        """Visibilities (complex128, channels x rows) of an image: NUFFTOperator.apply."""
        img = np.ascontiguousarray(image, dtype=np.float64).reshape(-1)
        vis = np.zeros(self.vis_space.size, dtype=np.complex128)
        B.check(B.lib().lt_nufft_forward_host(self._handle, B.ptr(img), B.ptr(vis)))
        return vis

    def adjoint(self, vis):
        """Image of N^T applied to visibilities (NUFFTOperator.T.apply)."""
        v = np.ascontiguousarray(vis, dtype=np.complex128).reshape(-1)
        img = np.zeros(self.image_space.size)
        B.check(B.lib().lt_nufft_adjoint_host(self._handle, B.ptr(v), B.ptr(img)))
        return img.reshape(self.image_space.shape)

    @property
    def has_weights(self):
        return getattr(self, '_has_weights', False)

    def set_weights(self, weights=None, sigma=None, mask=None):
        """Keep the per-visibility weights (or sigma^-2 on the unmasked
        visibilities, data_base.py:36-44) resident in the plan, together with
        the normal stencil built from them (FP64 plans, support <= 4); solvers
        created without weights use both."""
        w = _weights(weights, sigma, mask)
        if w.size != self.vis_space.size:
            raise ValueError(f'{w.size} weights for {self.vis_space.size} visibilities')
        B.check(B.lib().lt_nufft_set_weights(self._handle, B.ptr(w)))
        self._has_weights = True

    def set_stencil_shape(self, samples_per_group=0):
        """Forces the normal stencil's form (0: autotuned): 1, 2, 4, 8 samples
        per workgroup of the vector form (the same bits), or -1, -2, -4, -8
        groups of 8 samples per workgroup of the FP64 matrix-core form (the
        same bits among themselves, the vector form's to rounding)."""
        B.check(B.lib().lt_nufft_set_stencil_shape(self._handle, int(samples_per_group)))

    def set_degrid_form(self, form=1):
        """The degridder's tap evaluation: 1 (default) the window's 2W taps as
        polynomials of the offset from the cell centre, 0 the 2W+1 taps around
        the nearest grid point with the window picked by the offset's sign."""
        B.check(B.lib().lt_nufft_set_degrid_form(self._handle, int(form)))

    def set_row_form(self, form=1):
        """The band row transforms of rows padded to 1024: 1 (default) the
        radix-4 transform with its first and last passes in registers and
        its twiddles in LDS, 0 the 256-thread radix-4 transform staged in LDS
        (the same bits, FP64 and FP32)."""
        B.check(B.lib().lt_nufft_set_row_form(self._handle, int(form)))

    def set_col_form(self, form=1):
        """The band's column transforms when the padded grid is 1024 x 1024:
        1 (default) the row transforms' radix-4 passes over the image rows'
        part of each column, 0 rocFFT over whole columns (the same values to
        rounding)."""
        B.check(B.lib().lt_nufft_set_col_form(self._handle, int(form)))

    def time(self, direction, reps=20):
        """Mean device time (ms) of one degrid (0) or grid (1) launch."""
        ms = C.c_double(0)
        B.check(B.lib().lt_nufft_time(self._handle, int(direction), int(reps), C.byref(ms)))
        return ms.value

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and B._lib is not None:
            B._lib.lt_nufft_destroy(h)
            self._handle = None


class BatchedSourceInversion:
    """(L_s^T D^T C D L_s + lambda_c Cc^T Cc + lambda_t I) x_s = L_s^T D^T C d
    for every sample s of a LensBatch."""

    def __init__(self, lens_batch, nufft_plan, data=None, sigma=None, weights=None, mask=None, lambda_curvature=0.0,
                 lambda_tikhonov=0.0, operator='nufft', dirty_images=None):
        """data: one visibility vector shared by every sample; or
        dirty_images: D^T C d_s per sample (or one), e.g. from
        NUFFTPlan.adjoint(weights * d_s), when each sample has its own data."""
        self.lens_batch = lens_batch
        self.plan = nufft_plan
        if (data is None) == (dirty_images is None):
            raise ValueError('give either data or dirty_images')
        d = None if data is None else np.ascontiguousarray(data, dtype=np.complex128).reshape(-1)
        if weights is None and sigma is None:
            if not nufft_plan.has_weights:
                raise ValueError('give weights or sigma, or make them resident with NUFFTPlan.set_weights')
            self._w = None  # the plan's resident weights (and normal stencil)
        else:
            self._w = _weights(weights, sigma, mask)
        wptr = None if self._w is None else B.ptr(self._w)
        self._d = d
        mode = {'nufft': B.LT_OP_NUFFT, 'bcb': B.LT_OP_BCB, 'nufft_explicit': B.LT_OP_NUFFT_EXPLICIT}[operator]
        h = C.c_void_p()
        if d is not None:
            B.check(B.lib().lt_solver_create(lens_batch._handle, nufft_plan._handle, wptr, B.ptr(self._d),
                                             float(lambda_curvature), float(lambda_tikhonov), mode, C.byref(h)))
        elif isinstance(dirty_images, DeviceArray):
            npix = lens_batch.image_space.size
            nd = dirty_images.size // npix
            if self._w is not None:
                raise ValueError('device-resident dirty images take the plan\'s resident weights')
            if nd * npix != dirty_images.size or nd not in (1, lens_batch.nsamples):
                raise ValueError('dirty_images: one image-space array, or one per sample')
            self._dirty = dirty_images  # keeps the device memory alive
            B.check(B.lib().lt_solver_create_resident(lens_batch._handle, nufft_plan._handle, dirty_images.ptr, nd,
                                                      float(lambda_curvature), float(lambda_tikhonov), mode,
                                                      C.byref(h)))
        else:
            dl = dirty_images if isinstance(dirty_images, (list, tuple)) else [dirty_images]
            npix = lens_batch.image_space.size
            if len(dl) not in (1, lens_batch.nsamples) or any(np.size(a) != npix for a in dl):
                raise ValueError('dirty_images: one image-space array, or one per sample')
            self._dirty = np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1)
                                                               for a in dl]))
            B.check(B.lib().lt_solver_create_dirty(lens_batch._handle, nufft_plan._handle, wptr,
                                                   B.ptr(self._dirty), len(dl), float(lambda_curvature),
                                                   float(lambda_tikhonov), mode, C.byref(h)))
        self._handle = h
        self._sizes = [s.size for s in lens_batch.source_spaces]
        self._off = np.concatenate([[0], np.cumsum(self._sizes)])

    def _split(self, flat):
        return [flat[self._off[k]:self._off[k + 1]].reshape(self.lens_batch.source_spaces[k].shape)
                for k in range(len(self._sizes))]

    def rhs(self):
        out = np.zeros(self._off[-1])
        B.check(B.lib().lt_solver_rhs(self._handle, B.ptr(out)))
        return self._split(out)

    def set_graphs(self, on=True):
        """Run the FP64 CG's iteration blocks as captured HIP graphs (default)
        or as individual launches; the results are identical."""
        B.check(B.lib().lt_solver_set_graphs(self._handle, 1 if on else 0))

    def set_flexible_depth(self, depth=1):
        """Flexible CG keeps each new direction A-orthogonal to the last
        `depth` directions (1: the last one only)."""
        B.check(B.lib().lt_solver_set_flexible_depth(self._handle, int(depth)))

    def set_learned_iterations(self, n=0):
        """precond='learned': the network corrects M for the first n
        iterations only (0: no limit)."""
        B.check(B.lib().lt_solver_set_learned_iterations(self._handle, int(n)))

    def set_lens_preconditioner(self, inner_steps=3, max_row=64, c=0.0):
        B.check(B.lib().lt_solver_set_lens_pc(self._handle, int(inner_steps), int(max_row), float(c)))

    def lens_preconditioner_matrix(self):
        """The lens preconditioner's M = c L_s^T L_s + lambda Cc^T Cc over all
        samples' source pixels: (scipy CSR matrix, c)."""
        import scipy.sparse as sp
        nnz, c = C.c_long(0), C.c_double(0)
        B.check(B.lib().lt_solver_lens_pc_matrix(self._handle, C.byref(nnz), C.byref(c), None, None, None))
        n = int(sum(self._sizes))
        rp = np.zeros(n + 1, dtype=np.int32)
        ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        va = np.zeros(max(nnz.value, 1), dtype=np.float64)
        B.check(B.lib().lt_solver_lens_pc_matrix(self._handle, C.byref(nnz), None, B.ptr(rp), B.ptr(ci), B.ptr(va)))
        return sp.csr_matrix((va[:nnz.value], ci[:nnz.value], rp), shape=(n, n)), c.value

    def set_rhs(self, bs):
        """Replace the right-hand sides (per-sample arrays)."""
        b = np.ascontiguousarray(np.concatenate([np.asarray(v).reshape(-1) for v in bs]), dtype=np.float64)
        B.check(B.lib().lt_solver_set_rhs(self._handle, B.ptr(b)))

    def apply(self, xs, precision='f64'):
        """A x per sample (FP64, or the FP32 operator of the precision form)."""
        B.check(B.lib().lt_solver_set_precision(self._handle, {'f64': B.LT_F64, 'f32': B.LT_F32}[precision], 20))
        x = np.ascontiguousarray(np.concatenate([np.asarray(v).reshape(-1) for v in xs]), dtype=np.float64)
        y = np.zeros_like(x)
        B.check(B.lib().lt_solver_apply(self._handle, B.ptr(x), B.ptr(y)))
        return self._split(y)

    def diag(self):
        """Jacobi estimate of diag(A) per sample."""
        d = np.zeros(self._off[-1])
        B.check(B.lib().lt_solver_diag(self._handle, B.ptr(d)))
        return self._split(d)

    def set_preconditioner(self, net):
        """Network (learned.PCNet) for solve(precond='learned')."""
        self._net = net
        B.check(B.lib().lt_solver_set_pcnet(self._handle, net._handle if net is not None else None))

    def train_preconditioner(self, net, residuals, epochs=1, lr=1e-3, lr_final=1e-4, seed=0):
        """Train `net` with this inversion's operators on residual vectors
        (each a list of per-sample arrays, or a flat array of all samples);
        epochs=0 only evaluates.  Returns the mean energy loss per epoch."""
        self.set_preconditioner(net)
        R = np.ascontiguousarray(np.stack([np.concatenate([np.asarray(v).reshape(-1) for v in r])
                                           if isinstance(r, (list, tuple)) else np.asarray(r).reshape(-1)
                                           for r in residuals]), dtype=np.float64)
        hist = np.zeros(max(epochs, 1))
        B.check(B.lib().lt_solver_train_pcnet(self._handle, B.ptr(R), R.shape[0], int(epochs), float(lr),
                                              float(lr_final), int(seed), B.ptr(hist)))
        return hist

    def residuals(self, iterations, precond='none'):
        """True residuals b - A x_k of the solver (precond as in solve)
        stopped after each k in `iterations` (k = 0 gives b), as flat arrays
        of all samples."""
        b = np.concatenate([v.reshape(-1) for v in self.rhs()])
        out = []
        for k in iterations:
            if k == 0:
                out.append(b.copy())
                continue
            xs, _, _ = self.solve(rtol=1e-12, maxiter=int(k), precond=precond)
            ax = np.concatenate([v.reshape(-1) for v in self.apply(xs)])
            out.append(b - ax)
        return out

    def fallbacks(self):
        """Per sample, for the last learned solve: (iterations that took
        m = M^-1 r because the corrected direction was not a descent
        direction, whether the sample switched its network off and went on
        with M alone)."""
        c = np.zeros(len(self._sizes), dtype=np.int32)
        pl = np.zeros(len(self._sizes), dtype=np.int32)
        B.check(B.lib().lt_solver_pc_fallbacks(self._handle, B.ptr(c), B.ptr(pl)))
        return c, pl.astype(bool)

    def solve(self, rtol=1e-6, maxiter=5000, precond='none', precision='f64', refine_every=20):
        """Returns (sources, iterations, true relative residuals).
        precond: 'none' (plain CG), 'jacobi', 'learned' (flexible CG with
        the 'lens' preconditioner corrected by the network given to
        set_preconditioner) or 'lens' (flexible CG with
        a few inner CG steps on lambda Cc^T Cc + c L_s^T L_s, the reference's
        R + w L^T L preconditioner; see set_lens_preconditioner).  precision 'f32': the
        operator runs in FP32 inside the FP64 CG, with the FP64 true residual
        refreshed every `refine_every` iterations (the residuals returned are
        always FP64 true residuals)."""
        pc = {'none': B.LT_PC_NONE, 'jacobi': B.LT_PC_JACOBI, 'learned': B.LT_PC_LEARNED,
              'lens': B.LT_PC_LENS}[precond]
        B.check(B.lib().lt_solver_set_precision(self._handle, {'f64': B.LT_F64, 'f32': B.LT_F32}[precision],
                                                int(refine_every)))
        n = len(self._sizes)
        x = np.zeros(self._off[-1])
        its = np.zeros(n, dtype=np.int32)
        res = np.zeros(n)
        B.check(B.lib().lt_solver_solve(self._handle, float(rtol), int(maxiter), pc, B.ptr(x), B.ptr(its),
                                        B.ptr(res)))
        return self._split(x), its, res

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and B._lib is not None:
            B._lib.lt_solver_destroy(h)
            self._handle = None
