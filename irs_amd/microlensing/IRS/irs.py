"""Inverse ray shooting: an `IRS` class over lib_irs.so with the interface of
the IPM class (microlensing/IPM/ipm.py:7-473) and the IRS<T> defaults and
results of the reference's include/irs.cuh (light_loss 0.01, num_rays_y 100,
integer ray counts, num_rays_shot / num_rays_received, t_shoot_rays). The
reference exposes IRS only through its command-line program.
"""

import ctypes

import numpy as np

from ..IPM.ipm import IPM
from . import lib_irs


class IRS(IPM):
    def __init__(self, *args, **kwargs):
        if kwargs.get("is_double"):
            raise ValueError("the IRS library has a float interface only")
        if kwargs.get("kernel"):
            raise ValueError("the reference-faithful kernel exists for IPM only")
        kwargs.pop("kernel", None)
        super().__init__(*args, **kwargs)

    @staticmethod
    def _library(is_double):
        return lib_irs.lib

    def close(self):
        if getattr(self, "obj", None):
            self.lib.irs_free(self.obj)
            self.obj = None

    @property
    def kernel(self):
        return 0

    @kernel.setter
    def kernel(self, value):
        if value not in (None, 0):
            raise ValueError("the reference-faithful kernel exists for IPM only")

    @property
    def t_shoot_rays(self):
        return self.lib.get_t_shoot_rays(self.obj)

    @property
    def num_rays_shot(self):
        return self.lib.get_num_rays_shot(self.obj)

    @property
    def num_rays_received(self):
        return self.lib.get_num_rays_received(self.obj)

    def _copy_results(self):
        shape = (self.num_pixels_y2, self.num_pixels_y1)
        self.num_rays = np.ctypeslib.as_array(self.lib.get_pixels(self.obj), shape=shape).copy()
        # magnification = counts / mean rays per pixel (view_output.ipynb, cell 6)
        self.magnifications = self.num_rays / self.num_rays_y
        if self.write_parities:
            self.num_rays_minima = np.ctypeslib.as_array(self.lib.get_pixels_minima(self.obj), shape=shape).copy()
            self.num_rays_saddles = np.ctypeslib.as_array(self.lib.get_pixels_saddles(self.obj), shape=shape).copy()
            self.magnifications_minima = self.num_rays_minima / self.num_rays_y
            self.magnifications_saddles = self.num_rays_saddles / self.num_rays_y
        else:
            self.num_rays_minima = self.num_rays_saddles = None
            self.magnifications_minima = self.magnifications_saddles = None
        self.stars = np.ctypeslib.as_array(self.lib.get_stars(self.obj), shape=(self.num_stars, 3)).copy()

    def run(self):
        if not self.lib.run(self.obj, self.verbose):
            raise Exception("Error running IRS")
        self._copy_results()

    def derive(self):
        if not self.lib.irs_derive(self.obj, self.verbose):
            raise Exception("Error deriving IRS parameters")

    def prepare(self):
        if not self.lib.irs_prepare(self.obj, self.verbose):
            raise Exception("Error preparing IRS")

    def build(self):
        if not self.lib.irs_build(self.obj, self.verbose):
            raise Exception("Error building the IRS tree")
        self.stars = np.ctypeslib.as_array(self.lib.get_stars(self.obj), shape=(self.num_stars, 3)).copy()

    def shoot(self, row0=0, row1=-1):
        if not self.lib.irs_shoot(self.obj, self.verbose, int(row0), int(row1)):
            raise Exception("Error shooting IRS rays")

    def finish(self):
        if not self.lib.irs_finish(self.obj, self.verbose):
            raise Exception("Error finishing IRS")
        self._copy_results()

    def ray_positions(self, row):
        """Source-plane positions of one lattice row of rays and the pixel each was binned into."""
        n = self.num_ray_threads[0]
        out = np.zeros((n, 3))
        if not self.lib.irs_ray_positions(self.obj, int(row), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise Exception("Error computing ray positions")
        return out[:, 0] + 1j * out[:, 1], out[:, 2].astype(np.int64)

    def deflect(self, z, want_mu=False):
        raise NotImplementedError("use IPM.deflect; IRS exposes ray_positions")

    def save(self):
        if not self.lib.save(self.obj, self.verbose):
            raise Exception("Error saving IRS")
