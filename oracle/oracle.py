"""ctypes wrapper of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker. The product (irs_amd/) never imports it.

The oracle restates the weisluke/irs IPM/IRS algorithm in FP64 C++ with
OpenMP (oracle.cpp); this module adds numpy conveniences on top of it.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

MASS_FUNCTIONS = {"equal": 0, "uniform": 1, "salpeter": 2, "kroupa": 3, "optical_depth": 4}


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Lens(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("gamma", ctypes.c_double), ("theta", ctypes.c_double),
                ("kappastar", ctypes.c_double), ("rectangular", ctypes.c_int),
                ("corner_re", ctypes.c_double), ("corner_im", ctypes.c_double),
                ("approx", ctypes.c_int), ("taylor_smooth", ctypes.c_int)]


class Grid(ctypes.Structure):
    _fields_ = [("center_x", ctypes.c_double * 2), ("half_length_x", ctypes.c_double * 2),
                ("ray_half_sep", ctypes.c_double * 2), ("num_ray_threads", ctypes.c_int * 2),
                ("center_y", ctypes.c_double * 2), ("half_length_y", ctypes.c_double * 2),
                ("num_pixels", ctypes.c_int * 2)]


class Input(ctypes.Structure):
    _fields_ = [("kappa_tot", ctypes.c_double), ("shear", ctypes.c_double), ("kappa_star", ctypes.c_double),
                ("theta_star", ctypes.c_double), ("mass_function", ctypes.c_int), ("m_solar", ctypes.c_double),
                ("m_lower", ctypes.c_double), ("m_upper", ctypes.c_double), ("light_loss", ctypes.c_double),
                ("rectangular", ctypes.c_int), ("approx", ctypes.c_int), ("safety_scale", ctypes.c_double),
                ("center_y", ctypes.c_double * 2), ("half_length_y", ctypes.c_double * 2),
                ("num_pixels_y", ctypes.c_int * 2), ("num_rays_y", ctypes.c_int)]


class Derived(ctypes.Structure):
    _fields_ = [("m_lower", ctypes.c_double), ("m_upper", ctypes.c_double), ("mean_mass", ctypes.c_double),
                ("mean_mass2", ctypes.c_double), ("mean_mass2_ln_mass", ctypes.c_double),
                ("mu_ave", ctypes.c_double), ("num_rays_x", ctypes.c_double),
                ("ray_half_sep", ctypes.c_double * 2), ("half_length_x", ctypes.c_double * 2),
                ("num_ray_threads", ctypes.c_int * 2), ("center_x", ctypes.c_double * 2),
                ("num_stars", ctypes.c_int), ("corner", ctypes.c_double * 2),
                ("alpha_error", ctypes.c_double), ("taylor_smooth", ctypes.c_int),
                ("kappa_star", ctypes.c_double), ("theta_star", ctypes.c_double), ("rectangular", ctypes.c_int),
                ("covered", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i, v = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.orc_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.orc_star_uniforms.argtypes = [ctypes.c_uint32, ctypes.c_int64, P(d)]
        L.orc_mass_moments.argtypes = [i, d, d, d, P(d)]
        L.orc_mass_sample.argtypes = [i, d, d, d, d]
        L.orc_mass_sample.restype = d
        L.orc_derived_params.argtypes = [P(Input), i, P(Derived)]
        L.orc_derived_params_file.argtypes = [P(Input), i, i, i, d, d, d, P(d), P(Derived)]
        L.orc_rescaled_corner.argtypes = [i, i, d, d, d, i, d, d, P(d)]
        L.orc_root_half_length.argtypes = [i, i, d, d]
        L.orc_root_half_length.restype = d
        L.orc_generate_stars.argtypes = [i, i, d, d, i, d, d, d, ctypes.c_uint32, P(d)]
        L.orc_star_params.argtypes = [i, i, d, d, d, P(d), P(d)]
        L.orc_tree_build.argtypes = [P(d), i, d]
        L.orc_tree_build.restype = v
        L.orc_tree_free.argtypes = [v]
        L.orc_tree_levels.argtypes = [v]
        L.orc_tree_num_nodes.argtypes = [v]
        L.orc_tree_stars.argtypes = [v, P(d)]
        L.orc_tree_export.argtypes = [v, P(d), P(i), P(i)]
        L.orc_expansion_order.argtypes = [d, i, d, d]
        L.orc_tree_fmm.argtypes = [v, i]
        L.orc_tree_coeffs.argtypes = [v, i, P(d)]
        L.orc_deflect_tree.argtypes = [v, P(Lens), P(d), i, P(d), P(d), P(i)]
        L.orc_deflect_direct.argtypes = [P(d), i, P(Lens), P(d), i, P(d)]
        L.orc_triangle_alloc.argtypes = [P(d), d, i, i, P(ctypes.c_longlong), P(d), i]
        L.orc_ipm_rows.argtypes = [v, P(Lens), P(Grid), P(ctypes.c_longlong), i, i, P(d), P(d), P(d)]
        L.orc_ipm_window.argtypes = [v, P(Lens), P(Grid), P(ctypes.c_longlong), ctypes.c_longlong, i, i, i, i, P(d)]
        L.orc_irs_rows.argtypes = [v, P(Lens), P(Grid), P(ctypes.c_longlong), i, P(i), P(i), P(i)]
        L.orc_irs_rows.restype = ctypes.c_longlong
        L.orc_irs_bin.argtypes = [d, d, d, d, i, i]
        L.orc_irs_bin.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def star_uniforms(seed, i):
    u = np.zeros(4)
    lib().orc_star_uniforms(seed, i, _ptr(u))
    return u


def mass_moments(mf, m_lower, m_upper, m_solar=1.0):
    out = np.zeros(3)
    lib().orc_mass_moments(MASS_FUNCTIONS[mf], m_lower, m_upper, m_solar, _ptr(out))
    return out


def mass_sample(mf, p, m_lower, m_upper, m_solar=1.0):
    return lib().orc_mass_sample(MASS_FUNCTIONS[mf], p, m_lower, m_upper, m_solar)


def derived_params(kappa_tot, shear, kappa_star, theta_star=1.0, mass_function="equal", m_solar=1.0,
                   m_lower=0.01, m_upper=50.0, light_loss=0.001, rectangular=0, approx=1, safety_scale=1.37,
                   center_y=(0.0, 0.0), half_length_y=(5.0, 5.0), num_pixels_y=(1000, 1000), num_rays_y=1,
                   single=False):
    """Derived geometry (ipm.cuh:370-565); single=True evaluates it in float,
    as the reference's float build (IPM<float>) does."""
    inp = _input(kappa_tot, shear, kappa_star, theta_star, mass_function, m_solar, m_lower, m_upper, light_loss,
                 rectangular, approx, safety_scale, center_y, half_length_y, num_pixels_y, num_rays_y)
    out = Derived()
    lib().orc_derived_params(ctypes.byref(inp), int(single), ctypes.byref(out))
    return out


def _input(kappa_tot, shear, kappa_star, theta_star, mass_function, m_solar, m_lower, m_upper, light_loss, rectangular,
           approx, safety_scale, center_y, half_length_y, num_pixels_y, num_rays_y):
    return Input(kappa_tot, shear, kappa_star, theta_star, MASS_FUNCTIONS[mass_function], m_solar, m_lower, m_upper,
                 light_loss, int(rectangular), int(approx), safety_scale, (ctypes.c_double * 2)(*center_y),
                 (ctypes.c_double * 2)(*half_length_y), (ctypes.c_int * 2)(*num_pixels_y), num_rays_y)


def derived_params_file(stars, rectangular, corner, theta, kappa_tot, shear, light_loss=0.001, approx=1,
                        safety_scale=1.37, center_y=(0.0, 0.0), half_length_y=(5.0, 5.0), num_pixels_y=(1000, 1000),
                        num_rays_y=1, single=False):
    """Derived geometry when the stars come from a file (ipm.cuh:405-432, 507-529):
    the Derived record and whether the file's field covers the shooting region."""
    s = np.ascontiguousarray(stars, dtype=np.float64)
    inp = _input(kappa_tot, shear, 0.0, theta, "equal", 1.0, 1.0, 1.0, light_loss, rectangular, approx,
                 safety_scale, center_y, half_length_y, num_pixels_y, num_rays_y)
    out = Derived()
    rc = lib().orc_derived_params_file(ctypes.byref(inp), int(single), len(s), int(rectangular), corner[0], corner[1],
                                       theta, _ptr(s), ctypes.byref(out))
    return out, rc == 0


def rescaled_corner(rectangular, corner, theta, num_stars, mean_mass_actual, kappa_star, single=False):
    """The field corner after the stars are drawn (ipm.cuh:717-731) and the tree's
    root half-length (ipm.cuh:746-754): (corner, root_half_length)."""
    out = np.zeros(3)
    lib().orc_rescaled_corner(int(rectangular), int(single), corner[0], corner[1], theta, int(num_stars),
                              mean_mass_actual, kappa_star, _ptr(out))
    return (out[0], out[1]), out[2]


def root_half_length(rectangular, corner, single=False):
    """The tree's root half-length for a field corner (ipm.cuh:746-754)."""
    return lib().orc_root_half_length(int(rectangular), int(single), corner[0], corner[1])


def generate_stars(n, rectangular, corner, mass_function, m_lower, m_upper, m_solar, seed):
    out = np.zeros((n, 3))
    lib().orc_generate_stars(n, int(rectangular), corner[0], corner[1], MASS_FUNCTIONS[mass_function],
                             m_lower, m_upper, m_solar, seed, _ptr(out))
    return out


def star_params(stars, rectangular, corner, theta):
    out = np.zeros(6)
    s = np.ascontiguousarray(stars, dtype=np.float64)
    lib().orc_star_params(len(s), int(rectangular), corner[0], corner[1], theta, _ptr(s), _ptr(out))
    return dict(kappa_star=out[0], m_lower=out[1], m_upper=out[2], mean_mass=out[3], mean_mass2=out[4],
                mean_mass2_ln_mass=out[5])


def expansion_order(theta, levels, root_half_length, alpha_error):
    return lib().orc_expansion_order(theta, levels, root_half_length, alpha_error)


class Tree:
    """The reference's quadtree + FMM over a star field (tree_node.cuh, fmm.cuh)."""

    def __init__(self, stars, root_half_length):
        s = np.ascontiguousarray(stars, dtype=np.float64)
        self.h = lib().orc_tree_build(_ptr(s), len(s), root_half_length)
        self.num_stars = len(s)
        self.p = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().orc_tree_levels(self.h)

    @property
    def num_nodes(self):
        return lib().orc_tree_num_nodes(self.h)

    def stars(self):
        out = np.zeros((self.num_stars, 3))
        lib().orc_tree_stars(self.h, _ptr(out))
        return out

    def export(self):
        n = self.num_nodes
        geom = np.zeros((n, 3))
        info = np.zeros((n, 9), dtype=np.int32)
        lists = np.zeros((n, 40), dtype=np.int32)
        lib().orc_tree_export(self.h, _ptr(geom), _ptr(info, ctypes.c_int), _ptr(lists, ctypes.c_int))
        return geom, info, lists

    def fmm(self, p):
        if lib().orc_tree_fmm(self.h, p) != 0:
            raise ValueError("bad expansion order")
        self.p = p

    def coeffs(self, which="local"):
        out = np.zeros((self.num_nodes, self.p + 1, 2))
        lib().orc_tree_coeffs(self.h, 1 if which == "local" else 0, _ptr(out))
        return out[..., 0] + 1j * out[..., 1]

    def deflect(self, lens, z, want_mu=False):
        z = np.ascontiguousarray(np.column_stack([np.real(z), np.imag(z)]), dtype=np.float64)
        n = len(z)
        w = np.zeros((n, 2))
        mu = np.zeros(n) if want_mu else None
        leaf = np.zeros(n, dtype=np.int32)
        lib().orc_deflect_tree(self.h, ctypes.byref(lens), _ptr(z), n, _ptr(w),
                               _ptr(mu) if want_mu else None, _ptr(leaf, ctypes.c_int))
        wc = w[:, 0] + 1j * w[:, 1]
        return (wc, mu, leaf) if want_mu else (wc, leaf)

    def ipm_rows(self, lens, grid, rows, use_direct=False, parities=False):
        n1, n2 = grid.num_pixels[0], grid.num_pixels[1]
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        m = np.zeros((n2, n1))
        mn = np.zeros((n2, n1)) if parities else None
        sd = np.zeros((n2, n1)) if parities else None
        lib().orc_ipm_rows(self.h, ctypes.byref(lens), ctypes.byref(grid), _ptr(rows, ctypes.c_longlong), len(rows),
                           int(use_direct), _ptr(m), _ptr(mn) if parities else None, _ptr(sd) if parities else None)
        if parities:
            return mn + sd, mn, sd
        return m

    def ipm_window(self, lens, grid, records, x0, y0, width, height):
        """IPM of the cells named by window records (lattice row, first tile column,
        cell mask), kept to map columns [x0, x0 + width) and rows [y0, y0 + height)."""
        rec = np.ascontiguousarray(records, dtype=np.int64).reshape(-1, 3)
        win = np.zeros((height, width))
        lib().orc_ipm_window(self.h, ctypes.byref(lens), ctypes.byref(grid), _ptr(rec, ctypes.c_longlong), len(rec),
                             int(x0), int(y0), int(width), int(height), _ptr(win))
        return win

    def irs_rows(self, lens, grid, rows, parities=False):
        n1, n2 = grid.num_pixels[0], grid.num_pixels[1]
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        m = np.zeros((n2, n1), dtype=np.int32)
        mn = np.zeros((n2, n1), dtype=np.int32) if parities else None
        sd = np.zeros((n2, n1), dtype=np.int32) if parities else None
        P = ctypes.c_int
        got = lib().orc_irs_rows(self.h, ctypes.byref(lens), ctypes.byref(grid), _ptr(rows, ctypes.c_longlong),
                                 len(rows), _ptr(m, P), _ptr(mn, P) if parities else None,
                                 _ptr(sd, P) if parities else None)
        return (m, mn, sd, got) if parities else (m, got)


def deflect_direct(stars, lens, z):
    s = np.ascontiguousarray(stars, dtype=np.float64)
    zz = np.ascontiguousarray(np.column_stack([np.real(z), np.imag(z)]), dtype=np.float64)
    w = np.zeros((len(zz), 2))
    lib().orc_deflect_direct(_ptr(s), len(s), ctypes.byref(lens), _ptr(zz), len(zz), _ptr(w))
    return w[:, 0] + 1j * w[:, 1]


def triangle_alloc(tri, factor, n1, n2, cap=4096):
    t = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(6))
    idx = np.zeros(cap, dtype=np.int64)
    val = np.zeros(cap)
    n = lib().orc_triangle_alloc(_ptr(t), factor, n1, n2, _ptr(idx, ctypes.c_longlong), _ptr(val), cap)
    n = min(n, cap)
    return idx[:n], val[:n]


def irs_bin(w1, w2, h1, h2, n1, n2):
    return lib().orc_irs_bin(w1, w2, h1, h2, n1, n2)


def make_lens(kappa, gamma, theta, kappastar, corner, rectangular=0, approx=1, taylor_smooth=1):
    return Lens(kappa, gamma, theta, kappastar, int(rectangular), corner[0], corner[1], int(approx), taylor_smooth)


def make_grid(d, center_y, half_length_y, num_pixels):
    """Grid from a Derived record (or any object with the same fields)."""
    return Grid((ctypes.c_double * 2)(*d.center_x), (ctypes.c_double * 2)(*d.half_length_x),
                (ctypes.c_double * 2)(*d.ray_half_sep), (ctypes.c_int * 2)(*d.num_ray_threads),
                (ctypes.c_double * 2)(*center_y), (ctypes.c_double * 2)(*half_length_y),
                (ctypes.c_int * 2)(*num_pixels))
