"""ctypes handle on bin/lib_ipm_double.so (double ABI), as microlensing/IPM/lib_ipm_double.py loads it."""

import ctypes
import os

from .. import _binding

lib = _binding.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "bin", "lib_ipm_double.so"),
                    ctypes.c_double)
