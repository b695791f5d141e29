"""Inverse polygon mapping: the `IPM` class of the reference's
microlensing/IPM/ipm.py:7-473, with the same constructor arguments, validated
properties, run()/save(), result attributes and plots, over this engine's
lib_ipm.so / lib_ipm_double.so.

Additions: `expansion_order` (force the FMM order), `precision` (0 = FP64,
1 = mixed), `kernel` (0 = fast, 1 = reference-faithful FP64), `m2l_mode`
(0 = FP64 recurrences, 1 = MFMA with FP32 up to p = 20 and FP64 above,
2 = FP32 MFMA, 3 = FP64 MFMA) with `t_m2l`/`m2l_pairs`, `tile_rows` (lattice
rows per tile of the shooting kernel; 0 = chosen per launch), the staged
`prepare/build/shoot/finish` used by the benchmark, `deflect` for point
evaluations, and `close()`, which releases the engine object (the reference
never frees it).
"""

import ctypes

import numpy as np

from . import lib_ipm
from . import lib_ipm_double
from .. import _plotting


class IPM(object):
    def __init__(self, kappa_tot: float = None, shear: float = None, smooth_fraction: float = None,
                 kappa_star: float = None, theta_star: float = None, mass_function: str = None,
                 m_solar: float = None, m_lower: float = None, m_upper: float = None, light_loss: float = None,
                 rectangular: bool = None, approx: bool = None, safety_scale: float = None, starfile: str = None,
                 center_y1: float = None, center_y2: float = None, half_length_y1: float = None,
                 half_length_y2: float = None, num_pixels_y1: int = None, num_pixels_y2: int = None,
                 num_rays_y: int = None, random_seed: int = None, write_stars: bool = None,
                 write_maps: bool = None, write_parities: bool = None, write_histograms: bool = None,
                 outfile_prefix: str = None, verbose: int = 0, is_double: bool = False,
                 expansion_order: int = None, precision: int = None, kernel: int = None, m2l_mode: int = None,
                 tile_rows: int = None):
        """Arguments as in the reference (ipm.py:15-45); see the class docstring for the additions."""
        self.is_double = is_double
        self.lib = self._library(is_double)
        self.obj = self.lib.init()
        self.verbose = verbose

        self.kappa_tot = kappa_tot
        self.shear = shear

        if starfile is None:
            if smooth_fraction is not None and kappa_star is None:
                self.kappa_star = (1 - smooth_fraction) * self.kappa_tot
            self.kappa_star = kappa_star
            self.theta_star = theta_star
            self.mass_function = mass_function
            self.m_solar = m_solar
            self.m_lower = m_lower
            self.m_upper = m_upper
            if self.m_lower > self.m_upper:
                raise ValueError("m_lower must be <= m_upper")
            self.random_seed = random_seed

        self.light_loss = light_loss
        self.rectangular = rectangular
        self.approx = approx
        self.safety_scale = safety_scale
        self.starfile = starfile
        self.center_y1 = center_y1
        self.center_y2 = center_y2
        self.half_length_y1 = half_length_y1
        self.half_length_y2 = half_length_y2
        self.num_pixels_y1 = num_pixels_y1
        self.num_pixels_y2 = num_pixels_y2
        self.num_rays_y = num_rays_y
        self.write_stars = write_stars
        self.write_maps = write_maps
        self.write_parities = write_parities
        self.write_histograms = write_histograms
        self.outfile_prefix = outfile_prefix
        self.expansion_order = expansion_order
        self.precision = precision
        self.kernel = kernel
        self.m2l_mode = m2l_mode
        self.tile_rows = tile_rows

        self.magnifications = None
        self.magnifications_minima = None
        self.magnifications_saddles = None
        self.stars = None

    @staticmethod
    def _library(is_double):
        return lib_ipm_double.lib if is_double else lib_ipm.lib

    def close(self):
        if getattr(self, "obj", None):
            self.lib.ipm_free(self.obj)
            self.obj = None

    def __del__(self):
        self.close()

    # ------------------------------------------------------------------ fields
    @property
    def kappa_tot(self):
        return self.lib.get_kappa_tot(self.obj)

    @kappa_tot.setter
    def kappa_tot(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("kappa_tot must be >= 0")
            self.lib.set_kappa_tot(self.obj, value)

    @property
    def shear(self):
        return self.lib.get_shear(self.obj)

    @shear.setter
    def shear(self, value):
        if value is not None:
            self.lib.set_shear(self.obj, value)

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
            elif value > self.kappa_tot:
                raise ValueError("kappa_star must be <= kappa_tot")
            self.lib.set_kappa_star(self.obj, value)

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
        if value is not None:
            if value < 0:
                raise ValueError("theta_star must be >= 0")
            self.lib.set_theta_star(self.obj, value)

    @property
    def mass_function(self):
        return self.lib.get_mass_function(self.obj).decode("utf-8")

    @mass_function.setter
    def mass_function(self, value):
        if value is not None:
            if value.lower() not in ["equal", "uniform", "salpeter", "kroupa", "optical_depth"]:
                raise ValueError("mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth")
            self.lib.set_mass_function(self.obj, value.lower().encode("utf-8"))

    @property
    def m_solar(self):
        return self.lib.get_m_solar(self.obj)

    @m_solar.setter
    def m_solar(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("m_solar must be >= 0")
            self.lib.set_m_solar(self.obj, value)

    @property
    def m_lower(self):
        return self.lib.get_m_lower(self.obj)

    @m_lower.setter
    def m_lower(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("m_lower must be >= 0")
            self.lib.set_m_lower(self.obj, value)

    @property
    def m_upper(self):
        return self.lib.get_m_upper(self.obj)

    @m_upper.setter
    def m_upper(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("m_upper must be >= 0")
            self.lib.set_m_upper(self.obj, value)

    @property
    def light_loss(self):
        return self.lib.get_light_loss(self.obj)

    @light_loss.setter
    def light_loss(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("light_loss must be > 0")
            elif value > 0.01:
                raise ValueError("light_loss must be <= 0.01")
            self.lib.set_light_loss(self.obj, value)

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
        if value is not None:
            if value < 1.1:
                raise ValueError("safety_scale must be >= 1.1")
            self.lib.set_safety_scale(self.obj, value)

    @property
    def starfile(self):
        return self.lib.get_starfile(self.obj).decode("utf-8")

    @starfile.setter
    def starfile(self, value):
        if value is not None:
            self.lib.set_starfile(self.obj, value.encode("utf-8"))

    @property
    def center_y1(self):
        return self.lib.get_center_y1(self.obj)

    @center_y1.setter
    def center_y1(self, value):
        if value is not None:
            self.lib.set_center_y1(self.obj, value)

    @property
    def center_y2(self):
        return self.lib.get_center_y2(self.obj)

    @center_y2.setter
    def center_y2(self, value):
        if value is not None:
            self.lib.set_center_y2(self.obj, value)

    @property
    def center(self):
        return (self.center_y1, self.center_y2)

    @property
    def half_length_y1(self):
        return self.lib.get_half_length_y1(self.obj)

    @half_length_y1.setter
    def half_length_y1(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("half_length_y1 must be > 0")
            self.lib.set_half_length_y1(self.obj, value)

    @property
    def half_length_y2(self):
        return self.lib.get_half_length_y2(self.obj)

    @half_length_y2.setter
    def half_length_y2(self, value):
        if value is not None:
            if value < 0:
                raise ValueError("half_length_y2 must be > 0")
            self.lib.set_half_length_y2(self.obj, value)

    @property
    def half_length(self):
        return (self.half_length_y1, self.half_length_y2)

    @property
    def num_pixels_y1(self):
        return self.lib.get_num_pixels_y1(self.obj)

    @num_pixels_y1.setter
    def num_pixels_y1(self, value):
        if value is not None:
            if value < 1:
                raise ValueError("num_pixels_y1 must be >= 1")
            self.lib.set_num_pixels_y1(self.obj, value)

    @property
    def num_pixels_y2(self):
        return self.lib.get_num_pixels_y2(self.obj)

    @num_pixels_y2.setter
    def num_pixels_y2(self, value):
        if value is not None:
            if value < 1:
                raise ValueError("num_pixels_y2 must be >= 1")
            self.lib.set_num_pixels_y2(self.obj, value)

    @property
    def num_pixels(self):
        return (self.num_pixels_y1, self.num_pixels_y2)

    @property
    def pixel_scales(self):
        return (2 * self.half_length[0] / self.num_pixels[0], 2 * self.half_length[1] / self.num_pixels[1])

    @property
    def num_rays_y(self):
        return self.lib.get_num_rays_y(self.obj)

    @num_rays_y.setter
    def num_rays_y(self, value):
        if value is not None:
            if value < 1:
                raise ValueError("num_rays_y must be >= 1")
            self.lib.set_num_rays_y(self.obj, value)

    @property
    def random_seed(self):
        return self.lib.get_random_seed(self.obj)

    @random_seed.setter
    def random_seed(self, value):
        if value is not None:
            self.lib.set_random_seed(self.obj, value)

    @property
    def write_stars(self):
        return bool(self.lib.get_write_stars(self.obj))

    @write_stars.setter
    def write_stars(self, value):
        if value is not None:
            self.lib.set_write_stars(self.obj, int(value))

    @property
    def write_maps(self):
        return bool(self.lib.get_write_maps(self.obj))

    @write_maps.setter
    def write_maps(self, value):
        if value is not None:
            self.lib.set_write_maps(self.obj, int(value))

    @property
    def write_parities(self):
        return bool(self.lib.get_write_parities(self.obj))

    @write_parities.setter
    def write_parities(self, value):
        if value is not None:
            self.lib.set_write_parities(self.obj, int(value))

    @property
    def write_histograms(self):
        return bool(self.lib.get_write_histograms(self.obj))

    @write_histograms.setter
    def write_histograms(self, value):
        if value is not None:
            self.lib.set_write_histograms(self.obj, int(value))

    @property
    def outfile_prefix(self):
        return self.lib.get_outfile_prefix(self.obj).decode("utf-8")

    @outfile_prefix.setter
    def outfile_prefix(self, value):
        if value is not None:
            self.lib.set_outfile_prefix(self.obj, value.encode("utf-8"))

    @property
    def expansion_order(self):
        return self.lib.get_expansion_order(self.obj)

    @expansion_order.setter
    def expansion_order(self, value):
        if value is not None:
            if value < 0 or value > 31:
                raise ValueError("expansion_order must be 0 (automatic) or between 1 and 31")
            self.lib.set_expansion_order(self.obj, int(value))

    @property
    def precision(self):
        return self.lib.get_precision(self.obj)

    @precision.setter
    def precision(self, value):
        if value is not None:
            if value not in (0, 1):
                raise ValueError("precision must be 0 (fp64) or 1 (mixed)")
            self.lib.set_precision(self.obj, int(value))

    @property
    def m2l_mode(self):
        return self.lib.get_m2l_mode(self.obj)

    @m2l_mode.setter
    def m2l_mode(self, value):
        if value is not None:
            if value not in (0, 1, 2, 3):
                raise ValueError("m2l_mode must be 0 (direct FP64), 1 (MFMA), 2 (FP32 MFMA) or 3 (FP64 MFMA)")
            self.lib.set_m2l_mode(self.obj, int(value))

    @property
    def tile_rows(self):
        return self.lib.get_tile_rows(self.obj)

    @tile_rows.setter
    def tile_rows(self, value):
        if value is not None:
            if value < 0 or value > 65536:
                raise ValueError("tile_rows must be 0 (chosen per launch) or between 1 and 65536")
            self.lib.set_tile_rows(self.obj, int(value))

    @property
    def last_tile_rows(self):
        """Lattice rows per tile of the last shooting launch."""
        return self.lib.get_last_tile_rows(self.obj)

    @property
    def t_m2l(self):
        """Seconds in the MFMA M2L kernels of the last build."""
        return self.lib.get_t_m2l(self.obj)

    @property
    def m2l_pairs(self):
        return self.lib.get_m2l_pairs(self.obj)

    @property
    def m2l_fp32(self):
        return bool(self.lib.get_m2l_fp32(self.obj))

    @property
    def kernel(self):
        return self.lib.get_kernel(self.obj)

    @kernel.setter
    def kernel(self, value):
        if value is not None:
            if value not in (0, 1):
                raise ValueError("kernel must be 0 (fast) or 1 (reference)")
            self.lib.set_kernel(self.obj, int(value))

    # -------------------------------------------------------------- derived
    @property
    def num_stars(self):
        return self.lib.get_num_stars(self.obj)

    @property
    def corner(self):
        return (self.lib.get_corner_x1(self.obj), self.lib.get_corner_x2(self.obj))

    @property
    def tree_levels(self):
        return self.lib.get_tree_levels(self.obj)

    @property
    def alpha_error(self):
        return self.lib.get_alpha_error(self.obj)

    @property
    def num_ray_threads(self):
        return (self.lib.get_num_ray_threads_1(self.obj), self.lib.get_num_ray_threads_2(self.obj))

    @property
    def root_half_length(self):
        return self.lib.get_root_half_length(self.obj)

    @property
    def mean_near_stars(self):
        return self.lib.get_mean_near_stars(self.obj)

    @property
    def cells_shot(self):
        return self.lib.get_cells_shot(self.obj)

    @property
    def t_shoot_cells(self):
        return self.lib.get_t_shoot_cells(self.obj)

    # -------------------------------------------------------------- actions
    def _copy_results(self):
        shape = (self.num_pixels_y2, self.num_pixels_y1)
        self.magnifications = np.ctypeslib.as_array(self.lib.get_pixels(self.obj), shape=shape).copy()
        if self.write_parities:
            self.magnifications_minima = np.ctypeslib.as_array(self.lib.get_pixels_minima(self.obj), shape=shape).copy()
            self.magnifications_saddles = np.ctypeslib.as_array(self.lib.get_pixels_saddles(self.obj), shape=shape).copy()
        else:
            self.magnifications_minima = None
            self.magnifications_saddles = None
        self.stars = np.ctypeslib.as_array(self.lib.get_stars(self.obj), shape=(self.num_stars, 3)).copy()

    def run(self):
        if not self.lib.run(self.obj, self.verbose):
            raise Exception("Error running IPM")
        self._copy_results()

    def derive(self):
        """Validate and derive the geometry without touching the GPU: num_stars,
        the lattice and the image-plane region become readable."""
        if not self.lib.ipm_derive(self.obj, self.verbose):
            raise Exception("Error deriving IPM parameters")

    def prepare(self):
        if not self.lib.ipm_prepare(self.obj, self.verbose):
            raise Exception("Error preparing IPM")

    def build(self):
        if not self.lib.ipm_build(self.obj, self.verbose):
            raise Exception("Error building the IPM tree")
        self.stars = np.ctypeslib.as_array(self.lib.get_stars(self.obj), shape=(self.num_stars, 3)).copy()

    def shoot(self, row0=0, row1=-1):
        if not self.lib.ipm_shoot(self.obj, self.verbose, int(row0), int(row1)):
            raise Exception("Error shooting IPM cells")

    def finish(self):
        if not self.lib.ipm_finish(self.obj, self.verbose):
            raise Exception("Error finishing IPM")
        self._copy_results()

    def deflect(self, z, want_mu=False):
        """Lens equation w(z) through the built tree (FP64 output; mu if asked)."""
        z = np.asarray(z, dtype=np.complex128).ravel()
        zz = np.ascontiguousarray(np.column_stack([z.real, z.imag]))
        w = np.zeros((len(z), 2))
        mu = np.zeros(len(z)) if want_mu else None
        P = ctypes.POINTER(ctypes.c_double)
        ok = self.lib.ipm_deflect(self.obj, zz.ctypes.data_as(P), len(z), w.ctypes.data_as(P),
                                  mu.ctypes.data_as(P) if want_mu else None)
        if not ok:
            raise Exception("Error deflecting points")
        wc = w[:, 0] + 1j * w[:, 1]
        return (wc, mu) if want_mu else wc

    def window_cells(self, x0, y0, width, height, margin=1.0, capacity=1 << 22):
        """The lattice cells that can reach columns [x0, x0 + width) and rows
        [y0, y0 + height) of the map (row 0 at the top), widened by margin pixels:
        an (n, 3) int64 array of (lattice row, first column of a 63-cell tile,
        bit mask of the tile's cells). Needs build()."""
        while True:
            rec = np.zeros((capacity, 3), dtype=np.int64)
            n = self.lib.ipm_window_cells(self.obj, int(x0), int(y0), int(width), int(height), float(margin),
                                          rec.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), capacity)
            if n < 0:
                raise Exception("Error finding the cells of a window")
            if n <= capacity:
                return rec[:n].copy()
            capacity = int(n)

    def export_tree(self):
        n = self.lib.get_num_nodes(self.obj)
        geom = np.zeros((n, 3))
        info = np.zeros((n, 9), dtype=np.int32)
        lists = np.zeros((n, 40), dtype=np.int32)
        P = ctypes.POINTER
        if not self.lib.ipm_export_tree(self.obj, geom.ctypes.data_as(P(ctypes.c_double)),
                                        info.ctypes.data_as(P(ctypes.c_int)), lists.ctypes.data_as(P(ctypes.c_int))):
            raise Exception("Error exporting the tree")
        return geom, info, lists

    def export_coeffs(self, which="local"):
        n = self.lib.get_num_nodes(self.obj)
        p = self.expansion_order
        out = np.zeros((n, p + 1, 2))
        if not self.lib.ipm_export_coeffs(self.obj, 1 if which == "local" else 0,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise Exception("Error exporting coefficients")
        return out[..., 0] + 1j * out[..., 1]

    @property
    def magnitudes(self):
        return -2.5 * np.log10(self.magnifications / np.abs(self.mu_ave))

    @property
    def magnitudes_minima(self):
        if self.magnifications_minima is None:
            raise ValueError("magnifications_minima is None")
        return -2.5 * np.log10(self.magnifications_minima / np.abs(self.mu_ave))

    @property
    def magnitudes_saddles(self):
        if self.magnifications_saddles is None:
            raise ValueError("magnifications_saddles is None")
        return -2.5 * np.log10(self.magnifications_saddles / np.abs(self.mu_ave))

    def save(self):
        if not self.lib.save(self.obj, self.verbose):
            raise Exception("Error saving IPM")

    def plot_map(self, fig, ax, **kwargs):
        """The map in magnitudes relative to |mu_ave|, bright pixels dark; the
        default colour range spans the central 95% of the pixels."""
        dm = self.magnitudes
        _plotting.image(fig, ax, dm, _plotting.box(self.center, self.half_length, self.theta_star),
                        self.half_length[0] / self.half_length[1], "microlensing $\\Delta m$ (magnitudes)",
                        _plotting.Y_LABELS, kwargs,
                        {"vmin": np.percentile(dm, 2.5), "vmax": np.percentile(dm, 97.5), "cmap": "viridis_r"},
                        flip_colorbar=True)

    def plot_hist(self, fig, ax, bins=None, **kwargs):
        """Density of the pixels' magnitudes in bins of 0.01 mag, brightest at
        the right."""
        dm = self.magnitudes
        if bins is None:
            bins = np.arange(np.min(dm) - 0.01, np.max(dm) + 0.01, 0.01)
        _plotting.density(ax, dm, bins, "microlensing $\\Delta m$ (magnitudes)", "p($\\Delta m$)", kwargs,
                          flip_x=True)
