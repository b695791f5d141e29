"""Shared configurations and helpers for the tests."""

import ctypes

import numpy as np

# C1 of the benchmark definition, shrunk so the FP64 oracle shoots the whole
# map in about a second: kappa = kappa_star = gamma = 0.3, equal masses,
# circular field, half-length 10 theta_E, 100 x 100 pixels, 4 rays per pixel.
C1_SMALL = dict(kappa_tot=0.3, shear=0.3, kappa_star=0.3, theta_star=1.0, mass_function="equal", m_solar=1.0,
                m_lower=1.0, m_upper=1.0, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
                center_y1=0.0, center_y2=0.0, half_length_y1=10.0, half_length_y2=10.0, num_pixels_y1=100,
                num_pixels_y2=100, num_rays_y=4, random_seed=1, write_stars=False, write_maps=False,
                write_histograms=False)

# C2 shape (near-critical, smooth matter, Kroupa) shrunk for the oracle.
C2_SMALL = dict(kappa_tot=0.45, shear=0.5, kappa_star=0.36, theta_star=1.0, mass_function="kroupa", m_solar=1.0,
                m_lower=0.08, m_upper=1.5, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
                center_y1=0.0, center_y2=0.0, half_length_y1=2.0, half_length_y2=2.0, num_pixels_y1=64,
                num_pixels_y2=64, num_rays_y=4, random_seed=7, write_stars=False, write_maps=False,
                write_histograms=False)


# The FMM stress workload: kappa = kappa_star = 0.8, shear 0.1, no smooth
# matter, Salpeter 0.1-1 M_sun, 20000^2 pixels over half-length 50 (3.3e6 stars).
C3_FULL = dict(kappa_tot=0.8, shear=0.1, kappa_star=0.8, theta_star=1.0, mass_function="salpeter", m_solar=1.0,
               m_lower=0.1, m_upper=1.0, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
               center_y1=0.0, center_y2=0.0, half_length_y1=50.0, half_length_y2=50.0, num_pixels_y1=20000,
               num_pixels_y2=20000, num_rays_y=1, random_seed=3, write_stars=False, write_maps=False,
               write_histograms=False)

def write_starfile(O, path, params, seed):
    """Oracle-generated star field (Philox, FP64) written in the reference's
    .bin layout (star.cuh:498-534), so that runs of the float and the double
    library and the oracle all shoot through identical stars."""
    mf = params["mass_function"]
    d = O.derived_params(params["kappa_tot"], params["shear"], params["kappa_star"], params["theta_star"], mf,
                         params["m_solar"], params["m_lower"], params["m_upper"], params["light_loss"],
                         int(params["rectangular"]), int(params["approx"]), params["safety_scale"],
                         (params["center_y1"], params["center_y2"]),
                         (params["half_length_y1"], params["half_length_y2"]),
                         (params["num_pixels_y1"], params["num_pixels_y2"]), params["num_rays_y"])
    stars = O.generate_stars(d.num_stars, int(params["rectangular"]), d.corner, mf, d.m_lower, d.m_upper,
                             params["m_solar"], seed)
    mean_mass = stars[:, 2].mean()
    c = np.array(d.corner)
    if params["rectangular"]:
        c = np.sqrt(c / c[::-1]) * np.sqrt(np.pi * params["theta_star"] ** 2 * d.num_stars * mean_mass /
                                           (4 * params["kappa_star"]))
    else:
        c = np.array([np.sqrt(params["theta_star"] ** 2 * d.num_stars * mean_mass / params["kappa_star"]), 0.0])
    with open(path, "wb") as fh:
        np.array([d.num_stars, int(params["rectangular"])], dtype=np.int32).tofile(fh)
        np.array([c[0], c[1], params["theta_star"]], dtype=np.float64).tofile(fh)
        stars.astype(np.float64).tofile(fh)
    return stars


def from_starfile(params, path):
    kw = {k: v for k, v in params.items() if k not in ("kappa_star", "theta_star", "mass_function", "m_solar",
                                                         "m_lower", "m_upper", "random_seed", "rectangular")}
    kw["starfile"] = str(path)
    return kw


class Mirror:
    """The oracle's view of a finished engine run: same stars, same derived
    geometry, same lens parameters. `stars` defaults to the engine's own
    (tree-ordered) stars as returned by the ABI."""

    def __init__(self, O, m, p=None, stars=None):
        lib = m.lib
        obj = m.obj
        self.stars = np.array(m.stars if stars is None else stars, dtype=np.float64)
        self.p = m.expansion_order if p is None else p
        corner_full = (lib.get_corner_x1(obj), lib.get_corner_x2(obj))
        if not m.rectangular:
            corner_full = (corner_full[0], 0.0)
        self.lens = O.make_lens(m.kappa_tot, m.shear, m.theta_star, m.kappa_star, corner_full, m.rectangular,
                                m.approx, lib.get_taylor_smooth(obj))
        self.grid = O.Grid((ctypes.c_double * 2)(lib.get_center_x1(obj), lib.get_center_x2(obj)),
                           (ctypes.c_double * 2)(lib.get_half_length_x1(obj), lib.get_half_length_x2(obj)),
                           (ctypes.c_double * 2)(lib.get_ray_half_sep_1(obj), lib.get_ray_half_sep_2(obj)),
                           (ctypes.c_int * 2)(lib.get_num_ray_threads_1(obj), lib.get_num_ray_threads_2(obj)),
                           (ctypes.c_double * 2)(m.center_y1, m.center_y2),
                           (ctypes.c_double * 2)(m.half_length_y1, m.half_length_y2),
                           (ctypes.c_int * 2)(m.num_pixels_y1, m.num_pixels_y2))
        self.root_half_length = float(lib.get_root_half_length(obj))
        self.tree = O.Tree(self.stars, self.root_half_length)
        self.tree.fmm(self.p)

    @property
    def num_cells(self):
        return self.grid.num_ray_threads[0], self.grid.num_ray_threads[1]

    def ipm_map(self, rows=None, use_direct=False, parities=False):
        if rows is None:
            rows = np.arange(self.grid.num_ray_threads[1])
        return self.tree.ipm_rows(self.lens, self.grid, rows, use_direct=use_direct, parities=parities)

    def ipm_window(self, records, x0, y0, width, height):
        return self.tree.ipm_window(self.lens, self.grid, records, x0, y0, width, height)

    def irs_map(self, rows=None, parities=False):
        if rows is None:
            rows = np.arange(self.grid.num_ray_threads[1])
        return self.tree.irs_rows(self.lens, self.grid, rows, parities=parities)


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def random_points(m, n, seed=0):
    """Uniform points in the shooting region of a finished run."""
    lib, obj = m.lib, m.obj
    rng = np.random.default_rng(seed)
    cx, cy = lib.get_center_x1(obj), lib.get_center_x2(obj)
    hx, hy = lib.get_half_length_x1(obj), lib.get_half_length_x2(obj)
    return (cx + hx * rng.uniform(-1, 1, n)) + 1j * (cy + hy * rng.uniform(-1, 1, n))
