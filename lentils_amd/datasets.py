"""Radio data sets: the data side of the inversion.

Counterpart of lentils/common/data_base.py (Dataset, RadioDataset): the data
vector, its noise and mask, the diagonal covariance sigma^-2 on unmasked
visibilities, the NUFFT as the blurring operator, the dirty image
N^T C d, the dirty beam N_b^T C 1 and the blurred covariance operator (BCB)
N^T C N replaced by a convolution with the dirty beam.  The dirty beam and
the beam convolution run in liblentils_hip.so (NUFFT gridding + rocFFT).
"""
import ctypes as C

import numpy as np

from . import backend as B
from .fitsio import read_fits_image
from .operators import DiagonalOperator, NUFFTOperator, Operator
from .spaces import ImageSpace, VisibilitySpace
from .uvfits import load_uvfits


class Dataset:
    """data, sigma and mask on a space; covariance diag(sigma^-2) on the
    unmasked entries (data_base.py:15-60)."""

    def __init__(self, data=None, sigma=None, mask=None, space=None):
        self.space = space
        self.dtype = space.dtype
        self.data = np.array(data, dtype=self.dtype, order='C')
        self.sigma = np.array(sigma, dtype=np.float64, order='C')
        self.mask = np.array(mask, dtype=np.bool_, order='C')

    def __repr__(self):
        return f'{type(self).__name__} with space {self.space}'

    @property
    def blurring_operator(self):
        raise NotImplementedError

    @property
    def blurred_covariance_operator(self):
        raise NotImplementedError

    @property
    def covariance_operator(self):
        try:
            return self._covariance_op
        except AttributeError:
            diag = np.zeros_like(self.sigma, dtype=self.space.dtype)
            diag[self.mask] = self.sigma[self.mask] ** -2
            self._covariance_op = DiagonalOperator(self.space, diag)
            return self._covariance_op

    @property
    def shape(self):
        return self.space.shape[-2:]

    @property
    def num_channels(self):
        return self.space.shape[0]


class BlurredCovarianceOperator(Operator):
    """x -> zpad^T irfft2(rfft2(roll(beam)) / N * rfft2(zpad x)) on an image
    space (data_base.py:115-129), on the GPU; symmetric, so its transpose is
    itself."""

    def __init__(self, image_space, beam):
        if image_space.nx % 2 or image_space.ny % 2:
            raise ValueError('the blurred covariance operator needs even image sizes')
        self._beam = np.ascontiguousarray(beam, dtype=np.float64).reshape(2 * image_space.nx, 2 * image_space.ny)
        self._transpose = self
        super().__init__(image_space, image_space)

    def _matrixfree_forward(self, vec):
        sp = self.space_right
        x = np.ascontiguousarray(vec, dtype=np.float64).reshape(-1)
        n = x.size // (sp.nx * sp.ny)
        if n * sp.nx * sp.ny != x.size:
            raise ValueError(f'Incompatible shapes! {sp.shape}, {np.shape(vec)}')
        out = np.zeros_like(x)
        B.check(B.lib().lt_bcb_apply_host(sp, B.ptr(self._beam), n, B.ptr(x), B.ptr(out)))
        return out

    _matrixfree_transpose = _matrixfree_forward


def dirty_beam(vis_space, image_space, weights, kernel_support=4, kb_beta=None):
    """N_b^T C 1 on the beam space of image_space (data_base.py:163-172),
    with the reference's default NUFFT unless kernel_support / kb_beta are
    given.  Returns (beam of shape (2nx, 2ny), its peak)."""
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    if w.size != vis_space.size:
        raise ValueError(f'weights: {w.size} values for {vis_space.size} visibilities')
    beam = np.zeros((2 * image_space.nx, 2 * image_space.ny))
    peak = C.c_double(0.0)
    B.check(B.lib().lt_dirty_beam_host(vis_space, image_space, B.ptr(w), int(kernel_support),
                                       0.0 if kb_beta is None else float(kb_beta), B.ptr(beam), C.byref(peak)))
    return beam, peak.value


class RadioDataset(Dataset):
    """Interferometric data (data_base.py:64-172): uvfits loading (random
    groups, RR/LL combined), the NUFFT of the image space, dirty image and
    beam, and the blurred covariance operator."""

    def __init__(self, uvfits=None, dirty_image_fits=None, dirty_beam_fits=None, image_space=None,
                 combine_stokes=True, mfs=True):
        if uvfits is not None:
            uvspace, data, sigma, mask = load_uvfits(uvfits, combine_stokes)
        else:
            uvspace = VisibilitySpace()
            data, sigma, mask = [0.], [0.], [0]
        super().__init__(data=data, sigma=sigma, mask=mask, space=uvspace)
        if image_space is not None:
            self.image_space = image_space
            self.nufft_operator = NUFFTOperator(self.space, self.image_space)
        if dirty_image_fits is not None and dirty_beam_fits is not None and image_space is not None:
            # FITS images are (..., y, x); the first plane, transposed to x-major
            dimg = np.ascontiguousarray(read_fits_image(dirty_image_fits).T[:, :, 0], dtype=np.float64)
            dbeam = np.ascontiguousarray(read_fits_image(dirty_beam_fits).T[:, :, 0], dtype=np.float64)
            self._dirty_image = dimg.reshape(image_space.shape)
            self.beam_space = self._make_beam_space()
            self._dirty_beam = dbeam.reshape(self.beam_space.shape)

    @property
    def blurring_operator(self):
        return self.nufft_operator

    @property
    def blurred_covariance_operator(self):
        try:
            return self._fcf
        except AttributeError:
            self._fcf = BlurredCovarianceOperator(self.image_space, self.dirty_beam)
            return self._fcf

    @property
    def covariance_operator_dft(self):
        try:
            return self._covariance_op_dft
        except AttributeError:
            diag = np.zeros_like(self.sigma, dtype=self.space.dtype)
            diag[self.mask] = self.sigma[self.mask] ** -2
            self._covariance_op_dft = DiagonalOperator(self.space, diag, options='r2c')
            return self._covariance_op_dft

    @property
    def weights(self):
        """sigma^-2 on unmasked visibilities, 0 elsewhere (real)."""
        w = np.zeros(self.sigma.shape)
        w[self.mask] = self.sigma[self.mask] ** -2
        return w

    @property
    def dirty_image(self):
        try:
            return self._dirty_image
        except AttributeError:
            self._dirty_image = self.nufft_operator.T * self.covariance_operator * self.data
            return self._dirty_image

    def _make_beam_space(self):
        nx, ny = 2 * self.image_space.nx, 2 * self.image_space.ny
        rx = self.image_space.dx * self.image_space.nx
        ry = self.image_space.dx * self.image_space.ny
        return ImageSpace(shape=(nx, ny), bounds=[(-rx, rx), (-ry, ry)])

    @property
    def dirty_beam(self):
        try:
            return self._dirty_beam
        except AttributeError:
            self.beam_space = self._make_beam_space()
            beam, _ = dirty_beam(self.space, self.image_space, self.weights)
            self._dirty_beam = beam.reshape(self.beam_space.shape)
            return self._dirty_beam
