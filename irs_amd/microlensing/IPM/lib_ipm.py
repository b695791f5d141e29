"""ctypes handle on bin/lib_ipm.so (float ABI), as microlensing/IPM/lib_ipm.py:5 loads it."""

import ctypes
import os

from .. import _binding

lib = _binding.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "bin", "lib_ipm.so"),
                    ctypes.c_float)
