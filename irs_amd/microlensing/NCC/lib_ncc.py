"""ctypes handle on bin/lib_ncc.so, declared as microlensing/NCC/lib_ncc.py:5-62 declares it
(double values throughout)."""

import ctypes
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "bin", "lib_ncc.so")

_FIELDS = [
    ("infile_prefix", "s"), ("center_y1", "d"), ("center_y2", "d"), ("half_length_y1", "d"),
    ("half_length_y2", "d"), ("num_pixels_y1", "i"), ("num_pixels_y2", "i"), ("over_sample", "i"),
    ("write_maps", "i"), ("write_histograms", "i"), ("outfile_prefix", "s"),
]


def load(path=_PATH):
    if not os.path.exists(path):
        raise OSError(f"engine library {path} is missing; build it with irs_amd.build()")
    lib = ctypes.CDLL(path)
    v, d, i, s = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
    kinds = {"d": d, "i": i, "s": s}
    lib.init.argtypes = []
    lib.init.restype = v
    for name, kind in _FIELDS:
        getattr(lib, f"set_{name}").argtypes = [v, kinds[kind]]
        getattr(lib, f"set_{name}").restype = None
        getattr(lib, f"get_{name}").argtypes = [v]
        getattr(lib, f"get_{name}").restype = kinds[kind]
    lib.get_num_crossings.argtypes = [v]
    lib.get_num_crossings.restype = ctypes.POINTER(i)
    for name in ("get_t_ncc", "get_t_reduce"):
        getattr(lib, name).argtypes = [v]
        getattr(lib, name).restype = d
    lib.run.argtypes = [v, i]
    lib.run.restype = ctypes.c_bool
    lib.save.argtypes = [v, i]
    lib.save.restype = ctypes.c_bool
    lib.ncc_free.argtypes = [v]
    lib.ncc_free.restype = None
    return lib


lib = load()
