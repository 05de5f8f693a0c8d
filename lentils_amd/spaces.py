"""Spaces: the reference's ImageSpace / VisibilitySpace / FourierSpace.

Like lentils/common/space_base.py:17-136 these classes *are* the ctypes
structures passed by value to the backend, so a space can be handed straight
to the C ABI.  Shapes, dtypes, bounds/pixel conventions and the
FourierSpace geometry (du = 1/(ARCSEC * padded extent), gcx, half_nv) are the
reference's.  Masks are numpy arrays (x-major, shape (nx, ny)) or FITS files
read by the small reader in lentils_amd.fitsio.
"""
import copy
import ctypes as C

import numpy as np

from .backend import (c_bool_p, c_double_p, c_int_p, delaunay_space_ctype, fourier_space_ctype,
                      image_space_ctype, visibility_space_ctype)

ARCSEC_TO_RADIANS = 4.8481368111e-6  # lentils/backend/common.h:7


class Space:

    def __init__(self, shape=(0,), dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype

    def new_vector(self, value=0.0):
        return np.full(self.shape, value, dtype=self.dtype)

    def copy(self):
        return copy.deepcopy(self)

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def points(self):
        raise NotImplementedError

    def __repr__(self):
        return f'{type(self).__name__}{self.shape}'


class VisibilitySpace(Space, visibility_space_ctype):
    """(num_channels, num_stokes, num_rows) complex128; uvw in wavelengths."""

    def __init__(self, channels=(0.,), uvcoords=np.array([0., 0., 0.]), stokes=('I',)):
        self.channels = np.array(channels, dtype=np.float64, order='C').flatten()
        self._c_channels = self.channels.ctypes.data_as(c_double_p)
        self.num_channels = self.channels.shape[0]
        self.uvw = np.array(uvcoords, dtype=np.float64, order='C').reshape((-1, 3))
        self._c_uvw = self.uvw.ctypes.data_as(c_double_p)
        self.num_rows = self.uvw.shape[0]
        self.num_stokes = len(stokes)
        if self.num_stokes != 1 or stokes[0] != 'I':
            raise ValueError("Only stokes 'I' is allowed for now.")
        Space.__init__(self, shape=(self.num_channels, self.num_stokes, self.num_rows), dtype=np.complex128)

    @property
    def points(self):
        return self.uvw


class FourierSpace(Space, fourier_space_ctype):
    """Half-plane grid of rfft2 of an image space (space_base.py:62-84)."""

    def __init__(self, image_space, channels=(0.,), stokes=('I',)):
        self.nu, self.nv = image_space.nx, image_space.ny
        self.half_nv = self.nv // 2 + 1
        lx = ARCSEC_TO_RADIANS * (image_space.xmax - image_space.xmin)
        ly = ARCSEC_TO_RADIANS * (image_space.ymax - image_space.ymin)
        self.du = 1.0 / lx
        self.dv = 1.0 / ly
        self.gcx = ARCSEC_TO_RADIANS * 0.5 * (image_space.xmax + image_space.xmin)
        self.gcy = ARCSEC_TO_RADIANS * 0.5 * (image_space.ymax + image_space.ymin)
        self.channels = np.array(channels, dtype=np.float64, order='C').flatten()
        self._c_channels = self.channels.ctypes.data_as(c_double_p)
        self.num_channels = 1
        self.num_stokes = 1
        Space.__init__(self, shape=(self.num_channels, self.num_stokes, self.nu, self.half_nv), dtype=np.complex128)


class ImageSpace(Space, image_space_ctype):
    """(1, 1, nx, ny) float64 pixel grid; pixel (i, j) has corner
    (xmin + i dx, ymin + j dy); rows are x-major (index = ny*i + j)."""

    def __init__(self, shape=(128, 128), bounds=((-1.0, 1.0), (-1.0, 1.0)), channels=(0.,), stokes=('I',),
                 mask=None):
        self.nx, self.ny = int(shape[0]), int(shape[1])
        self.num_channels = 1
        self.num_stokes = 1
        self.bounds = np.array(bounds, dtype=np.float64, order='C').flatten()
        self.xmin, self.xmax = self.bounds[0], self.bounds[1]
        self.ymin, self.ymax = self.bounds[2], self.bounds[3]
        self.dx = (self.xmax - self.xmin) / self.nx
        self.dy = (self.ymax - self.ymin) / self.ny
        self.channels = np.array(channels, dtype=np.float64, order='C').flatten()
        self._c_channels = self.channels.ctypes.data_as(c_double_p)
        if mask is None:
            self.mask = np.ones((self.nx, self.ny), dtype=np.bool_)
        elif isinstance(mask, str):
            from .fitsio import read_fits_image
            self.mask = np.ascontiguousarray(read_fits_image(mask)[:, :].T.astype(np.bool_))
        else:
            self.mask = np.ascontiguousarray(np.asarray(mask).reshape(self.nx, self.ny).astype(np.bool_))
        self._c_mask = self.mask.ctypes.data_as(c_bool_p)
        Space.__init__(self, shape=(self.num_channels, self.num_stokes, self.nx, self.ny), dtype=np.float64)

    @property
    def points(self):
        """Pixel corner coordinates, (nx, ny, 2) (space_base.py:128-136)."""
        pts = np.zeros((self.nx, self.ny, 2))
        pts[:, :, 0] = (self.xmin + self.dx * np.arange(self.nx))[:, None]
        pts[:, :, 1] = (self.ymin + self.dy * np.arange(self.ny))[None, :]
        return pts

    def _refresh(self):
        """Re-point the C fields after the mask array was replaced."""
        self.mask = np.ascontiguousarray(self.mask, dtype=np.bool_)
        self._c_mask = self.mask.ctypes.data_as(c_bool_p)


class DelaunaySpace(Space, delaunay_space_ctype):
    """A triangulated source plane (space_base.py:139-158): the Delaunay
    triangulation of the given points (scipy.spatial.Delaunay, as the
    reference), its convex-hull vertices (`edge`), and the ctypes fields of
    delaunay_space; vectors have shape (1, 1, num_points)."""

    def __init__(self, points, channels=(0.,), stokes=('I',)):
        from scipy.spatial import Delaunay
        self.num_channels = 1
        self.num_stokes = 1
        self.delaunay = Delaunay(points)
        self.edge = np.unique(self.delaunay.convex_hull.flatten())
        self.vertices = np.ascontiguousarray(self.delaunay.points, dtype=np.float64)
        self._c_points = self.vertices.ctypes.data_as(c_double_p)
        self.num_points = self.vertices.shape[0]
        self.triangles = np.ascontiguousarray(self.delaunay.simplices, dtype=np.int32)
        self._c_triangles = self.triangles.ctypes.data_as(c_int_p)
        self.num_triangles = self.triangles.shape[0]
        Space.__init__(self, shape=(self.num_channels, self.num_stokes, self.num_points), dtype=np.float64)

    @property
    def points(self):
        return self.vertices
