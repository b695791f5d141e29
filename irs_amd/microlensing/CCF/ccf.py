"""Critical curves and caustics of a star field: the CCF class of
microlensing/CCF/ccf.py:6-361 over lib_ccf.so, with the same constructor
arguments, property validation, run() / save() and plots."""

import numpy as np

from . import lib_ccf
from .. import _plotting

MASS_FUNCTIONS = ("equal", "uniform", "salpeter", "kroupa", "optical_depth")


class CCF(object):
    def __init__(self, kappa_tot: float = None, shear: float = None, smooth_fraction: float = None, kappa_star: float = None,
                 theta_star: float = None, mass_function: str = None, m_solar: float = None, m_lower: float = None,
                 m_upper: float = None, rectangular: bool = None, approx: bool = None, safety_scale: float = None,
                 num_stars: int = None, starfile: str = None, num_phi: int = None, num_branches: int = None,
                 random_seed: int = None, write_stars: bool = None, write_critical_curves: bool = None,
                 write_caustics: bool = None, write_length_scales: bool = None, outfile_prefix: str = None,
                 verbose: int = 0):
        """Critical curves and caustics of a star field (the reference's CCF
        class, ccf.py:7-78; same argument names, meanings and checks).

        Lens: kappa_tot and shear describe the macro-model; the stellar part is
        kappa_star, or smooth_fraction (the share of kappa_tot that is smooth
        matter) when kappa_star is not given. theta_star is the Einstein radius
        of a one-solar-mass lens and m_solar the solar mass, both in the caller's
        units.

        Stars: mass_function is one of equal, uniform, Salpeter, Kroupa or
        optical_depth, cut to [m_lower, m_upper] solar masses. The field is a
        rectangle or a disk (rectangular), safety_scale times as large as the
        region it must cover, with num_stars stars drawn from random_seed (0 is
        kept for fields read from a file). A starfile, binary (count, shape,
        corner, theta_star, then the stars) or text (x1 x2 m per line with
        theta_star = 1), replaces every star setting. approx selects the
        approximate terms of the smooth-matter deflection.

        Curves: num_phi steps of the phase over [0, 2 pi], split into
        num_branches branches (num_phi even and a multiple of 2 * num_branches).

        Output: write_stars, write_critical_curves, write_caustics and
        write_length_scales choose the files written under outfile_prefix.
        """
        self.lib = lib_ccf.lib
        self.obj = self.lib.init()
        self.verbose = verbose

        self.kappa_tot = kappa_tot
        self.shear = shear

        if starfile is None:
            if smooth_fraction is not None and kappa_star is None:
                kappa_star = (1 - smooth_fraction) * self.kappa_tot
            self.kappa_star = kappa_star
            self.theta_star = theta_star
            self.mass_function = mass_function
            self.m_solar = m_solar
            self.m_lower = m_lower
            self.m_upper = m_upper
            if self.m_lower > self.m_upper:
                raise ValueError("m_lower must be <= m_upper")
            self.num_stars = num_stars
            self.random_seed = random_seed

        self.rectangular = rectangular
        self.approx = approx
        self.safety_scale = safety_scale

        self.starfile = starfile

        self.num_phi = num_phi
        if self.num_phi % 2 != 0:
            raise ValueError("num_phi must be an even number")
        self.num_branches = num_branches
        if self.num_phi % (2 * self.num_branches) != 0:
            raise ValueError("num_phi must be a multiple of 2 * num_branches")

        self.write_stars = write_stars
        self.write_critical_curves = write_critical_curves
        self.write_caustics = write_caustics
        self.write_mu_length_scales = write_length_scales

        self.outfile_prefix = outfile_prefix

    def close(self):
        if self.obj:
            self.lib.ccf_free(self.obj)
            self.obj = None

    def _num(self, name, value, minimum=None, maximum=None, cast=float):
        if value is None:
            return
        if minimum is not None and value < minimum:
            raise ValueError(f"{name} must be >= {minimum}")
        if maximum is not None and value > maximum:
            raise ValueError(f"{name} must be <= {maximum}")
        getattr(self.lib, f"set_{name}")(self.obj, cast(value))

    @property
    def kappa_tot(self):
        return self.lib.get_kappa_tot(self.obj)

    @kappa_tot.setter
    def kappa_tot(self, value):
        self._num("kappa_tot", value, 0)

    @property
    def shear(self):
        return self.lib.get_shear(self.obj)

    @shear.setter
    def shear(self, value):
        self._num("shear", value)

    @property
    def mu_ave(self):
        return 1 / ((1 - self.kappa_tot) ** 2 - self.shear ** 2)

    @property
    def kappa_star(self):
        return self.lib.get_kappa_star(self.obj)

    @kappa_star.setter
    def kappa_star(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("kappa_star must be >= 0")
            if value > self.kappa_tot:
                raise ValueError("kappa_star must be <= kappa_tot")
            self.lib.set_kappa_star(self.obj, float(value))

    @property
    def stellar_fraction(self):
        return self.kappa_star / self.kappa_tot

    @property
    def smooth_fraction(self):
        return 1 - self.kappa_star / self.kappa_tot

    @property
    def theta_star(self):
        return self.lib.get_theta_star(self.obj)

    @theta_star.setter
    def theta_star(self, value):
        self._num("theta_star", value, 0)

    @property
    def mass_function(self):
        return self.lib.get_mass_function(self.obj).decode("utf-8")

    @mass_function.setter
    def mass_function(self, value):
        if value is not None:
            if value.lower() not in MASS_FUNCTIONS:
                raise ValueError("mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth")
            self.lib.set_mass_function(self.obj, value.lower().encode("utf-8"))

    @property
    def m_solar(self):
        return self.lib.get_m_solar(self.obj)

    @m_solar.setter
    def m_solar(self, value):
        self._num("m_solar", value, 0)

    @property
    def m_lower(self):
        return self.lib.get_m_lower(self.obj)

    @m_lower.setter
    def m_lower(self, value):
        self._num("m_lower", value, 0)

    @property
    def m_upper(self):
        return self.lib.get_m_upper(self.obj)

    @m_upper.setter
    def m_upper(self, value):
        self._num("m_upper", value, 0)

    @property
    def rectangular(self):
        return bool(self.lib.get_rectangular(self.obj))

    @rectangular.setter
    def rectangular(self, value):
        if value is not None:
            self.lib.set_rectangular(self.obj, int(value))

    @property
    def approx(self):
        return bool(self.lib.get_approx(self.obj))

    @approx.setter
    def approx(self, value):
        if value is not None:
            self.lib.set_approx(self.obj, int(value))

    @property
    def safety_scale(self):
        return self.lib.get_safety_scale(self.obj)

    @safety_scale.setter
    def safety_scale(self, value):
        self._num("safety_scale", value, 1.1)

    @property
    def num_stars(self):
        return self.lib.get_num_stars(self.obj)

    @num_stars.setter
    def num_stars(self, value):
        self._num("num_stars", value, 1, cast=int)

    @property
    def starfile(self):
        return self.lib.get_starfile(self.obj).decode("utf-8")

    @starfile.setter
    def starfile(self, value):
        if value is not None:
            self.lib.set_starfile(self.obj, value.encode("utf-8"))

    @property
    def num_phi(self):
        return self.lib.get_num_phi(self.obj)

    @num_phi.setter
    def num_phi(self, value):
        self._num("num_phi", value, 1, cast=int)

    @property
    def num_branches(self):
        return self.lib.get_num_branches(self.obj)

    @num_branches.setter
    def num_branches(self, value):
        self._num("num_branches", value, 1, cast=int)

    @property
    def random_seed(self):
        return self.lib.get_random_seed(self.obj)

    @random_seed.setter
    def random_seed(self, value):
        if value is not None:
            self.lib.set_random_seed(self.obj, int(value))

    def _flag(name):
        def get(self):
            return bool(getattr(self.lib, f"get_{name}")(self.obj))

        def put(self, value):
            if value is not None:
                getattr(self.lib, f"set_{name}")(self.obj, int(value))
        return property(get, put)

    write_stars = _flag("write_stars")
    write_critical_curves = _flag("write_critical_curves")
    write_caustics = _flag("write_caustics")
    write_mu_length_scales = _flag("write_mu_length_scales")

    @property
    def outfile_prefix(self):
        return self.lib.get_outfile_prefix(self.obj).decode("utf-8")

    @outfile_prefix.setter
    def outfile_prefix(self, value):
        if value is not None:
            self.lib.set_outfile_prefix(self.obj, value.encode("utf-8"))

    @property
    def corner(self):
        return (self.lib.get_corner_x1(self.obj), self.lib.get_corner_x2(self.obj))

    @property
    def num_roots(self):
        return self.lib.get_num_roots(self.obj)

    def _copy(self, getter, shape, wanted=True):
        """A copy of one of the library's result arrays, or None when that
        output was not asked for."""
        if not wanted:
            return None
        return np.array(np.ctypeslib.as_array(getattr(self.lib, getter)(self.obj), shape=shape))

    def run(self):
        """Find the critical curves (and, as asked, the caustics and the
        magnification length scales). Curves are (roots x branches) rows of
        num_phi / num_branches + 1 points, as (re, im) pairs."""
        if not self.lib.run(self.obj, self.verbose):
            raise Exception("Error running CCF")
        curves = (self.num_roots * self.num_branches, self.num_phi // self.num_branches + 1)
        self.critical_curves = self._copy("get_critical_curves", curves + (2,))
        self.caustics = self._copy("get_caustics", curves + (2,), self.write_caustics)
        self.mu_length_scales = self._copy("get_mu_length_scales", curves, self.write_mu_length_scales)
        self.stars = self._copy("get_stars", (self.num_stars, 3))

    @property
    def t_ccs(self):
        return self.lib.get_t_ccs(self.obj)

    @property
    def t_init_roots(self):
        return self.lib.get_t_init_roots(self.obj)

    @property
    def max_error(self):
        return self.lib.get_max_error(self.obj)

    def save(self):
        if not self.lib.save(self.obj, self.verbose):
            raise Exception("Error saving CCF")

    def plot_critical_curves(self, fig, ax, **kwargs):
        _plotting.curves(ax, self.critical_curves, _plotting.X_LABELS, kwargs)

    def plot_caustics(self, fig, ax, **kwargs):
        _plotting.curves(ax, self.caustics, _plotting.Y_LABELS, kwargs)
