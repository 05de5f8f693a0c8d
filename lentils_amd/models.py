"""Lens models: GlobalLensModel, PowerLawEllipsoid, ExternalPotential.

Parameter packing into the 320-byte generic_mass_model records follows
lentils/models/lens_models.py:102-116 (PEMD) and :129-148 (shear);
setup_raytracing sorts components by z_l and drops z_l >= z_s
(lens_models.py:27-59).  The angular-diameter distances the reference fills
from astropy are never read by the deflection code (deflect.c uses only
fpars), so they are left zero here.  Deflection runs on the GPU
(lt_deflect_points).
"""
import ctypes as C

import numpy as np

from . import backend as B
from .backend import generic_mass_model_ctype, global_lens_model_ctype


class GlobalLensModel(global_lens_model_ctype):

    def __init__(self, components=None, cosmology=None):
        self.cosmology = cosmology
        self.components = []
        for comp in components or []:
            self.add_component(comp)
        self.lenses = np.zeros(0, dtype=generic_mass_model_ctype)

    def add_component(self, component):
        self.components.append(component)

    def setup_raytracing(self, z_s):
        lenses = np.zeros(len(self.components), dtype=generic_mass_model_ctype)
        for i, comp in enumerate(self.components):
            lenses[i] = comp.get_cpars()
        lenses.sort(order='z_l', kind='stable')
        lenses = lenses[lenses['z_l'] < z_s]
        lenses['z_s'] = z_s
        self.lenses = np.ascontiguousarray(lenses)
        self.num_lenses = len(self.lenses)
        self.num_zplanes = 1
        self._c_lenses = self.lenses.ctypes.data_as(C.c_void_p)

    def deflect(self, points, z_s, deriv=False):
        """Source-plane positions of image-plane points (..., 2).

        deriv=True returns (deflected, gradients) as lens_models.py:64-77
        does: gradients has shape (8,) + points.shape and, through the
        reference's deflect_points(want_deriv=1), receives in its first 12
        values the derivatives d alpha / d(b_eff, qh, f, x, y, th) of every
        PEMD component summed over all points (deflect.c:304-338); the
        per-point values are deflect_gradients()."""
        self.setup_raytracing(z_s)
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros_like(pts)
        n = pts.size // 2
        if deriv:
            gradients = np.zeros((8,) + pts.shape)
            B.lib().deflect_points(self, B.ptr(pts), n, B.ptr(out), 1, B.ptr(gradients))
            return out, gradients
        B.check(B.lib().lt_deflect_points(C.byref(self), B.ptr(pts), n, B.ptr(out)))
        return out

    def deflect_gradients(self, points, z_s):
        """(deflected, dalpha): dalpha has shape (n_pemd, 6) + points.shape,
        the derivatives of each PEMD component's deflection with respect to
        its (b_eff, qh, f, x, y, th) at every point (b_eff the internal
        Einstein-radius parameter of deflect.c:284, held fixed for the other
        five; th in degrees)."""
        self.setup_raytracing(z_s)
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros_like(pts)
        n = pts.size // 2
        npemd = int(np.sum(self.lenses['type'] == 0))
        dv = np.zeros((n, max(npemd, 1), 12))
        B.check(B.lib().lt_deflect_points_deriv(C.byref(self), B.ptr(pts), n, B.ptr(out), B.ptr(dv)))
        dv = dv[:, :npemd].reshape(n, npemd, 6, 2).transpose(1, 2, 0, 3)
        return out, np.ascontiguousarray(dv).reshape((npemd, 6) + pts.shape)


class MassModel:

    def get_cpars(self):
        raise NotImplementedError


class PowerLawEllipsoid(MassModel):
    """PEMD: Einstein-radius-like b, axis ratio f, position angle th (deg),
    slope qh (gamma = 2 qh + 1), centre (x, y); rc is carried but unused, as
    in the reference."""
    _my_ctype = 0

    def __init__(self, z=0.881000, b=0.463544, th=-14.278754, f=0.799362, x=-0.046847, y=-0.105357, rc=0.000571,
                 qh=0.506730):
        self.z, self.x, self.y, self.f, self.th, self.b, self.qh, self.rc = z, x, y, f, th, b, qh, rc

    def get_cpars(self):
        c = np.zeros(1, dtype=generic_mass_model_ctype)[0]
        c['type'] = self._my_ctype
        c['z_l'] = self.z
        c['fpars'][:9] = [self.x, self.y, self.f, self.th, np.sin((self.th + 90) * np.pi / 180),
                          np.cos((self.th + 90) * np.pi / 180), self.b, self.qh, self.rc]
        return c


class ExternalPotential(MassModel):
    """External shear (ss, sa in deg) plus Keeton-2001 convergence / shear
    gradient terms (gks, gka, gss, gsa)."""
    _my_ctype = 1

    def __init__(self, z=0.881000, x=-0.046847, y=-0.105357, ss=-0.046500, sa=7.921300, gks=0.0, gka=0.0, gss=0.0,
                 gsa=0.0):
        self.z, self.x, self.y, self.ss, self.sa = z, x, y, ss, sa
        self.gks, self.gka, self.gss, self.gsa = gks, gka, gss, gsa

    def get_cpars(self):
        c = np.zeros(1, dtype=generic_mass_model_ctype)[0]
        c['type'] = self._my_ctype
        c['z_l'] = self.z
        f = c['fpars']
        f[0], f[1], f[2], f[3] = self.x, self.y, self.ss, self.sa
        f[4], f[5] = np.sin((self.sa + 90) * np.pi / 180), np.cos((self.sa + 90) * np.pi / 180)
        f[6], f[7] = self.gks, self.gka
        f[8], f[9] = np.sin((self.gka + 90) * np.pi / 180), np.cos((self.gka + 90) * np.pi / 180)
        f[10], f[11] = self.gss, self.gsa
        f[12], f[13] = np.sin((self.gsa + 90) * np.pi / 180), np.cos((self.gsa + 90) * np.pi / 180)
        return c
