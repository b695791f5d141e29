"""ctypes handle on bin/lib_ccf.so, declared as microlensing/CCF/lib_ccf.py:5-116 declares it
(double values throughout)."""

import ctypes
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "bin", "lib_ccf.so")

_FIELDS = [
    ("kappa_tot", "d"), ("shear", "d"), ("kappa_star", "d"), ("theta_star", "d"), ("mass_function", "s"),
    ("m_solar", "d"), ("m_lower", "d"), ("m_upper", "d"), ("rectangular", "i"), ("approx", "i"),
    ("safety_scale", "d"), ("num_stars", "i"), ("starfile", "s"), ("num_phi", "i"), ("num_branches", "i"),
    ("random_seed", "i"), ("write_stars", "i"), ("write_critical_curves", "i"), ("write_caustics", "i"),
    ("write_mu_length_scales", "i"), ("outfile_prefix", "s"),
]


def load(path=_PATH):
    if not os.path.exists(path):
        raise OSError(f"engine library {path} is missing; build it with irs_amd.build()")
    lib = ctypes.CDLL(path)
    v, d, i, s, P = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER
    kinds = {"d": d, "i": i, "s": s}
    lib.init.argtypes = []
    lib.init.restype = v
    for name, kind in _FIELDS:
        getattr(lib, f"set_{name}").argtypes = [v, kinds[kind]]
        getattr(lib, f"set_{name}").restype = None
        getattr(lib, f"get_{name}").argtypes = [v]
        getattr(lib, f"get_{name}").restype = kinds[kind]
    for name, res in (("get_num_roots", i), ("get_corner_x1", d), ("get_corner_x2", d), ("get_stars", P(d)),
                      ("get_critical_curves", P(d)), ("get_caustics", P(d)), ("get_mu_length_scales", P(d)),
                      ("get_t_ccs", d), ("get_t_init_roots", d), ("get_max_error", d),
                      ("get_expansion_order", i), ("get_tree_levels", i)):
        getattr(lib, name).argtypes = [v]
        getattr(lib, name).restype = res
    lib.run.argtypes = [v, i]
    lib.run.restype = ctypes.c_bool
    lib.save.argtypes = [v, i]
    lib.save.restype = ctypes.c_bool
    lib.ccf_free.argtypes = [v]
    lib.ccf_free.restype = None
    return lib


lib = load()
