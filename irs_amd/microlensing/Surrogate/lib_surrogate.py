"""ctypes signatures of lib_surrogate.so (include/lib_surrogate.h)."""

import ctypes
import os

from ... import BIN_DIR

SYMBOLS = ["surrogate_create", "surrogate_free", "surrogate_num_params", "surrogate_set_params", "surrogate_get_params",
           "surrogate_load_data", "surrogate_train_step", "surrogate_gradients", "surrogate_predict",
           "surrogate_time_predict"]


def load(path=None):
    path = path or os.path.join(BIN_DIR, "lib_surrogate.so")
    if not os.path.exists(path):
        raise OSError(f"engine library {path} is missing; build it with irs_amd.build()")
    lib = ctypes.CDLL(path)
    v, i, u, b, d, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_bool, ctypes.c_double, ctypes.c_longlong
    F = ctypes.POINTER(ctypes.c_float)
    I = ctypes.POINTER(ctypes.c_int)

    def sig(name, args, res):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res

    sig("surrogate_create", [i, u], v)
    sig("surrogate_free", [v], None)
    sig("surrogate_num_params", [v], ll)
    sig("surrogate_set_params", [v, F], b)
    sig("surrogate_get_params", [v, F], b)
    sig("surrogate_load_data", [v, F, F, i], b)
    sig("surrogate_train_step", [v, I, i, ctypes.c_float, ctypes.c_float], d)
    sig("surrogate_gradients", [v, F, F, i, F], d)
    sig("surrogate_predict", [v, F, i, F], b)
    sig("surrogate_time_predict", [v, F, i, i], d)
    return lib


lib = load()
