"""ctypes handle on bin/lib_irs.so (float ABI)."""

import ctypes
import os

from .. import _binding

lib = _binding.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "bin", "lib_irs.so"),
                    ctypes.c_float, irs=True)
