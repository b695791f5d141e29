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

def write_starfile(O, path, params, seed, margin=1.0):
    """Oracle-generated star field (Philox, FP64) written in the reference's
    .bin layout (star.cuh:498-534), so that runs of the float and the double
    library and the oracle all shoot through identical stars. margin > 1 draws
    margin^2 times the stars over a field margin times wider (the same
    kappa_star): the region a star file must cover grows with the realised
    <m^2> / <m> (ipm.cuh:405-432, 521-529), which for a few thousand stars can
    exceed the mass function's."""
    mf = params["mass_function"]
    d = O.derived_params(params["kappa_tot"], params["shear"], params["kappa_star"], params["theta_star"], mf,
                         params["m_solar"], params["m_lower"], params["m_upper"], params["light_loss"],
                         int(params["rectangular"]), int(params["approx"]), params["safety_scale"],
                         (params["center_y1"], params["center_y2"]),
                         (params["half_length_y1"], params["half_length_y2"]),
                         (params["num_pixels_y1"], params["num_pixels_y2"]), params["num_rays_y"])
    n = int(np.ceil(d.num_stars * margin * margin)) if margin != 1.0 else d.num_stars
    corner = tuple(margin * c for c in d.corner)
    stars = O.generate_stars(n, int(params["rectangular"]), corner, mf, d.m_lower, d.m_upper, params["m_solar"], seed)
    # the file's field is the one the stars were drawn in, so it covers the
    # shooting region (ipm.cuh:521-529) whatever the realised mean mass; its
    # kappa_star is that of the realisation (ipm.cuh:405-432)
    c = np.array(corner) if params["rectangular"] else np.array([np.hypot(*corner), 0.0])
    with open(path, "wb") as fh:
        np.array([n, int(params["rectangular"])], dtype=np.int32).tofile(fh)
        np.array([c[0], c[1], params["theta_star"]], dtype=np.float64).tofile(fh)
        stars.astype(np.float64).tofile(fh)
    return stars


def from_starfile(params, path):
    kw = {k: v for k, v in params.items() if k not in ("kappa_star", "theta_star", "mass_function", "m_solar",
                                                         "m_lower", "m_upper", "random_seed", "rectangular")}
    kw["starfile"] = str(path)
    return kw


def read_starfile(path):
    """A star file in the reference's .bin layout (star.cuh:175-320), float or
    double, told apart by its size: (n, rectangular, corner, theta, stars AoS)."""
    import os
    size = os.path.getsize(path)
    with open(path, "rb") as fh:
        n, rect = (int(x) for x in np.fromfile(fh, dtype=np.int32, count=2))
        dt = np.float64 if size == 8 + 8 * 3 + 24 * n else np.float32
        head = np.fromfile(fh, dtype=dt, count=3).astype(np.float64)
        stars = np.fromfile(fh, dtype=dt, count=3 * n).astype(np.float64).reshape(n, 3)
    return n, rect, (float(head[0]), float(head[1])), float(head[2]), stars


def oracle_geometry(O, m):
    """The oracle's derived parameters for an engine object's inputs: from its
    star file when it has one (ipm.cuh:405-432), else from the mass function
    (ipm.cuh:379-399), evaluated in the library's value type (float for
    lib_ipm.so and lib_irs.so, as the reference's float build). The engine
    scales m_lower and m_upper by m_solar in place (ipm.cuh:381-390), so the
    inputs are recovered by the inverse scaling."""
    y = dict(light_loss=m.light_loss, approx=int(m.approx), safety_scale=m.safety_scale,
             center_y=(m.center_y1, m.center_y2), half_length_y=(m.half_length_y1, m.half_length_y2),
             num_pixels_y=(m.num_pixels_y1, m.num_pixels_y2), num_rays_y=m.num_rays_y, single=not m.is_double)
    if m.starfile:
        _, rect, corner, theta, stars = read_starfile(m.starfile)
        d, covered = O.derived_params_file(stars, rect, corner, theta, m.kappa_tot, m.shear, **y)
        assert covered
        return d, (rect, corner, theta, stars)
    mf = m.mass_function.lower()
    ms = m.m_solar
    ml, mu = (1.0, 1.0) if mf == "equal" else (m.m_lower / ms, m.m_upper / ms)
    d = O.derived_params(m.kappa_tot, m.shear, m.kappa_star, m.theta_star, mf, ms, ml, mu,
                         rectangular=int(m.rectangular), **y)
    return d, None


class Mirror:
    """The oracle's view of a finished engine run.

    The oracle derives the geometry itself (oracle_geometry): the image-plane
    region, the lattice, the star count, the smooth-term order, the star moments
    of the realised field (star.cuh:124-161), the field corner re-scaled with
    them (ipm.cuh:717-731) and the tree's root half-length (ipm.cuh:746-754).
    Every one of the engine's derived getters is checked against it: the
    integers exactly, the values to 1e-12 relative for the FP64 library and to
    4e-7 (three float ulps) for a float library, whose derivation the oracle
    evaluates in float as the reference's float build does. The moments of a
    field drawn on the device are sums in another order than the oracle's:
    1e-10 in FP64. The oracle then shoots with its own FP64 geometry; a float
    library's lattice, after that check, is the engine's float one, the
    lattice that build shoots.

    `stars` defaults to the engine's own (tree-ordered) stars from the ABI."""

    def __init__(self, O, m, p=None, stars=None):
        lib = m.lib
        obj = m.obj
        self.stars = np.array(m.stars if stars is None else stars, dtype=np.float64)
        self.p = m.expansion_order if p is None else p
        d, file = oracle_geometry(O, m)
        self.derived = d
        fp64 = bool(m.is_double)
        single = not fp64
        rtol = 1e-12 if fp64 else 4e-7

        def close(got, want, what, tol=None):
            np.testing.assert_allclose(got, want, rtol=rtol if tol is None else tol, atol=0, err_msg=what)

        assert lib.get_num_stars(obj) == d.num_stars == len(self.stars)
        assert (lib.get_num_ray_threads_1(obj), lib.get_num_ray_threads_2(obj)) == tuple(d.num_ray_threads)
        assert lib.get_taylor_smooth(obj) == d.taylor_smooth
        close([lib.get_center_x1(obj), lib.get_center_x2(obj)], list(d.center_x), "center_x")
        close([lib.get_half_length_x1(obj), lib.get_half_length_x2(obj)], list(d.half_length_x), "half_length_x")
        close([lib.get_ray_half_sep_1(obj), lib.get_ray_half_sep_2(obj)], list(d.ray_half_sep), "ray_half_sep")
        close(lib.get_alpha_error(obj), d.alpha_error, "alpha_error")
        close(m.kappa_star, d.kappa_star, "kappa_star")
        # moments of the realised field, with the corner the reference takes
        # them over: the derived one, or the file's
        rect = d.rectangular
        sp_corner = file[1] if file else tuple(d.corner)
        engine_stars = np.array(m.stars, dtype=np.float64)
        sp = O.star_params(engine_stars, rect, sp_corner, d.theta_star)
        close([lib.get_kappa_star_actual(obj), lib.get_mean_mass_actual(obj), lib.get_mean_mass2_actual(obj),
               lib.get_mean_mass2_ln_mass_actual(obj), lib.get_m_lower_actual(obj), lib.get_m_upper_actual(obj)],
              [sp["kappa_star"], sp["mean_mass"], sp["mean_mass2"], sp["mean_mass2_ln_mass"], sp["m_lower"],
               sp["m_upper"]], "star moments", tol=1e-10 if fp64 else 4e-7)
        if file:
            corner = tuple(d.corner)
            root = O.root_half_length(rect, corner, single)
        else:
            mm = float(np.float32(sp["mean_mass"])) if single else sp["mean_mass"]
            corner, root = O.rescaled_corner(rect, tuple(d.corner), d.theta_star, d.num_stars, mm, d.kappa_star,
                                             single)
        want = corner if rect else (float(np.hypot(*corner)), 0.0)
        # a device-drawn field's corner and root scale with its realised mean mass
        grown = 1e-10 if (fp64 and not file) else None
        close([lib.get_corner_x1(obj), lib.get_corner_x2(obj)], list(want), "corner", tol=grown)
        close(lib.get_root_half_length(obj), root, "root_half_length", tol=grown)
        self.root_half_length = root if fp64 else float(lib.get_root_half_length(obj))
        corner_full = want if fp64 else ((lib.get_corner_x1(obj), lib.get_corner_x2(obj)) if rect
                                         else (lib.get_corner_x1(obj), 0.0))
        self.lens = O.make_lens(m.kappa_tot, m.shear, d.theta_star, d.kappa_star if fp64 else m.kappa_star, corner_full,
                                rect, m.approx, d.taylor_smooth)
        if fp64:
            geo = (d.center_x, d.half_length_x, d.ray_half_sep)
        else:
            geo = ((lib.get_center_x1(obj), lib.get_center_x2(obj)),
                   (lib.get_half_length_x1(obj), lib.get_half_length_x2(obj)),
                   (lib.get_ray_half_sep_1(obj), lib.get_ray_half_sep_2(obj)))
        D2, I2 = ctypes.c_double * 2, ctypes.c_int * 2
        self.grid = O.Grid(D2(*geo[0]), D2(*geo[1]), D2(*geo[2]), I2(*d.num_ray_threads),
                           D2(m.center_y1, m.center_y2), D2(m.half_length_y1, m.half_length_y2),
                           I2(m.num_pixels_y1, m.num_pixels_y2))
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
