"""ctypes signatures of the engine's C ABI (include/lib_ipm.h, include/lib_irs.h).

The reference declares the same IPM signatures by hand in
microlensing/IPM/lib_ipm.py:7-135 and lib_ipm_double.py; here one table
serves the float and the double library and the IRS library.
"""

import ctypes
import os

_FIELDS = [
    # (name, kind) with kind "d" (dtype), "i" (int), "s" (string)
    ("kappa_tot", "d"), ("shear", "d"), ("kappa_star", "d"), ("theta_star", "d"), ("mass_function", "s"),
    ("m_solar", "d"), ("m_lower", "d"), ("m_upper", "d"), ("light_loss", "d"), ("rectangular", "i"),
    ("approx", "i"), ("safety_scale", "d"), ("starfile", "s"), ("center_y1", "d"), ("center_y2", "d"),
    ("half_length_y1", "d"), ("half_length_y2", "d"), ("num_pixels_y1", "i"), ("num_pixels_y2", "i"),
    ("num_rays_y", "i"), ("random_seed", "i"), ("write_stars", "i"), ("write_maps", "i"),
    ("write_parities", "i"), ("write_histograms", "i"), ("outfile_prefix", "s"),
    ("expansion_order", "i"), ("precision", "i"), ("m2l_mode", "i"), ("tile_rows", "i"),
]


_DERIVED = ["get_center_x1", "get_center_x2", "get_half_length_x1", "get_half_length_x2", "get_ray_half_sep_1",
            "get_ray_half_sep_2", "get_kappa_star_actual", "get_mean_mass", "get_mean_mass2", "get_mean_mass_actual",
            "get_mu_ave", "get_mean_mass2_actual", "get_mean_mass2_ln_mass", "get_mean_mass2_ln_mass_actual",
            "get_m_lower_actual", "get_m_upper_actual"]


def load(path, dtype, irs=False):
    """Load one engine library and declare every entry point it exports."""
    if not os.path.exists(path):
        raise OSError(f"engine library {path} is missing; build it with irs_amd.build()")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    v, i, d, s, b, ll = ctypes.c_void_p, ctypes.c_int, dtype, ctypes.c_char_p, ctypes.c_bool, ctypes.c_longlong
    kinds = {"d": d, "i": i, "s": s}
    lib.init.argtypes = []
    lib.init.restype = v
    for name, kind in _FIELDS:
        getattr(lib, f"set_{name}").argtypes = [v, kinds[kind]]
        getattr(lib, f"set_{name}").restype = None
        getattr(lib, f"get_{name}").argtypes = [v]
        getattr(lib, f"get_{name}").restype = kinds[kind]

    def sig(name, args, res):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res

    sig("get_num_stars", [v], i)
    sig("get_corner_x1", [v], d)
    sig("get_corner_x2", [v], d)
    sig("get_stars", [v], P(d))
    sig("get_tree_levels", [v], i)
    sig("get_alpha_error", [v], d)
    sig("get_num_ray_threads_1", [v], i)
    sig("get_num_ray_threads_2", [v], i)
    sig("get_root_half_length", [v], d)
    sig("get_mean_near_stars", [v], ctypes.c_double)
    sig("get_num_nodes", [v], i)
    sig("get_cells_shot", [v], ll)
    sig("get_taylor_smooth", [v], i)
    sig("get_t_m2l", [v], ctypes.c_double)
    sig("get_m2l_pairs", [v], ll)
    sig("get_m2l_fp32", [v], i)
    sig("get_last_tile_rows", [v], i)
    for name in _DERIVED:
        sig(name, [v], d)
    sig("run", [v, i], b)
    sig("save", [v, i], b)
    if irs:
        sig("irs_free", [v], None)
        sig("get_pixels", [v], P(ctypes.c_int))
        sig("get_pixels_minima", [v], P(ctypes.c_int))
        sig("get_pixels_saddles", [v], P(ctypes.c_int))
        sig("get_t_shoot_rays", [v], ctypes.c_double)
        sig("get_num_rays_shot", [v], ll)
        sig("get_num_rays_received", [v], ll)
        sig("irs_prepare", [v, i], b)
        sig("irs_derive", [v, i], b)
        sig("irs_build", [v, i], b)
        sig("irs_shoot", [v, i, ll, ll], b)
        sig("irs_finish", [v, i], b)
        sig("irs_ray_positions", [v, ll, P(ctypes.c_double)], b)
    else:
        sig("ipm_free", [v], None)
        sig("set_kernel", [v, i], None)
        sig("get_kernel", [v], i)
        sig("get_pixels", [v], P(d))
        sig("get_pixels_minima", [v], P(d))
        sig("get_pixels_saddles", [v], P(d))
        sig("get_t_shoot_cells", [v], ctypes.c_double)
        sig("ipm_prepare", [v, i], b)
        sig("ipm_build", [v, i], b)
        sig("ipm_shoot", [v, i, ll, ll], b)
        sig("ipm_finish", [v, i], b)
        sig("ipm_deflect", [v, P(ctypes.c_double), i, P(ctypes.c_double), P(ctypes.c_double)], b)
        sig("ipm_deflect_noalloc", [v, P(ctypes.c_double), i, P(ctypes.c_double), P(ctypes.c_double)], b)
        sig("ipm_window_cells", [v, i, i, i, i, ctypes.c_double, P(ll), ll], ll)
        sig("ipm_export_tree", [v, P(ctypes.c_double), P(i), P(i)], b)
        sig("ipm_export_coeffs", [v, i, P(ctypes.c_double)], b)
        sig("ipm_derive", [v, i], b)
        sig("ipm_batch_create", [], v)
        sig("ipm_batch_free", [v], None)
        sig("ipm_batch_add", [v, v], i)
        sig("ipm_batch_shoot", [v, i], b)
        sig("get_batch_t_shoot", [v], ctypes.c_double)
        sig("get_batch_cells", [v], ll)
        sig("get_batch_size", [v], i)
        sig("ipm_batch_histograms", [v, P(ctypes.c_uint)], b)
        sig("get_batch_tile_h", [v], i)
        sig("get_batch_blocks_per_cu", [v], i)
        sig("get_batch_t_tune", [v], ctypes.c_double)
        sig("ipm_batch_set_tuning", [v, i, i], b)
    return lib


_M2L_SYMBOLS = ["set_m2l_mode", "get_m2l_mode", "get_t_m2l", "get_m2l_pairs", "get_m2l_fp32", "set_tile_rows",
                "get_tile_rows", "get_last_tile_rows"]

# every symbol include/lib_ipm.h declares (the first 63 are the reference's)
IPM_SYMBOLS = (["init"] + [f"set_{n}" for n, _ in _FIELDS[:26]] + [f"get_{n}" for n, _ in _FIELDS[:26]] +
               ["get_num_stars", "get_corner_x1", "get_corner_x2", "get_stars", "get_pixels", "get_pixels_minima",
                "get_pixels_saddles", "get_t_shoot_cells", "run", "save"] +
               ["ipm_free", "set_expansion_order", "get_expansion_order", "set_precision", "get_precision",
                "set_kernel", "get_kernel", "get_tree_levels", "get_alpha_error", "get_num_ray_threads_1",
                "get_num_ray_threads_2", "get_root_half_length", "get_mean_near_stars", "get_taylor_smooth"] + _DERIVED +
               ["ipm_prepare", "ipm_build",
                "ipm_shoot", "ipm_finish", "get_cells_shot", "ipm_deflect", "ipm_deflect_noalloc", "ipm_window_cells", "get_num_nodes", "ipm_export_tree",
                "ipm_export_coeffs", "ipm_derive", "ipm_batch_create", "ipm_batch_free", "ipm_batch_add",
                "ipm_batch_shoot", "get_batch_t_shoot", "get_batch_cells", "get_batch_size", "ipm_batch_histograms",
                "get_batch_tile_h", "get_batch_blocks_per_cu", "get_batch_t_tune", "ipm_batch_set_tuning"] + _M2L_SYMBOLS)

IRS_SYMBOLS = (["init", "irs_free"] + [f"set_{n}" for n, _ in _FIELDS[:26]] + [f"get_{n}" for n, _ in _FIELDS[:26]] +
               ["get_num_stars", "get_corner_x1", "get_corner_x2", "get_stars", "get_pixels", "get_pixels_minima",
                "get_pixels_saddles", "get_t_shoot_rays", "get_num_rays_shot", "get_num_rays_received", "run", "save",
                "set_expansion_order", "get_expansion_order", "set_precision", "get_precision", "get_tree_levels",
                "get_alpha_error", "get_num_ray_threads_1", "get_num_ray_threads_2", "get_root_half_length",
                "get_mean_near_stars", "get_num_nodes", "get_cells_shot", "get_taylor_smooth"] + _DERIVED +
               ["irs_prepare", "irs_derive", "irs_build", "irs_shoot",
                "irs_finish", "irs_ray_positions"] + _M2L_SYMBOLS)
