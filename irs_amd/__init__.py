"""MI355X-native microlensing magnification-map engine (IPM and IRS).

The compiled engine lives in irs_amd/bin (lib_ipm.so, lib_ipm_double.so,
lib_irs.so, lib_surrogate.so, lib_ccf.so, lib_ncc.so), built from irs_amd/csrc by
`build()`. The Python interface mirrors the reference's `microlensing`
package: see irs_amd.microlensing.
"""

import os
import subprocess

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
BIN_DIR = os.path.join(PACKAGE_DIR, "bin")
CSRC_DIR = os.path.join(PACKAGE_DIR, "csrc")
LIBRARIES = ("lib_ipm.so", "lib_ipm_double.so", "lib_irs.so", "lib_surrogate.so", "lib_ccf.so", "lib_ncc.so")


def library_path(name):
    return os.path.join(BIN_DIR, name)


def build(jobs=8, force=False):
    """Compile the HIP engine for gfx950 with the Makefile in csrc/."""
    cmd = ["make", "-C", CSRC_DIR, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


def built():
    return all(os.path.exists(library_path(n)) for n in LIBRARIES)
