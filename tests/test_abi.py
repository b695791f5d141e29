"""The C ABI (include/lib_ipm.h, include/lib_irs.h) without a GPU: every
declared symbol is exported, setters and getters round-trip, parameter
validation fails with the reference's messages before any device call
(ipm.cuh:235-368, star.cuh:175-320), and the Python mirror keeps the
reference's ValueError checks (microlensing/IPM/ipm.py:94-395)."""

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "irs_amd", "bin")


def header_symbols(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\((?:IPMHandle|IRSHandle|IPMBatch|SurrogateHandle|CCFHandle|NCCHandle|void|int "
                                 r"max_batch)", text)))


@pytest.mark.parametrize("lib,header,least", [("lib_ipm.so", "lib_ipm.h", 60), ("lib_ipm_double.so", "lib_ipm.h", 60),
                                              ("lib_irs.so", "lib_irs.h", 60), ("lib_surrogate.so", "lib_surrogate.h", 10),
                                              ("lib_ccf.so", "lib_ccf.h", 53), ("lib_ncc.so", "lib_ncc.h", 26)])
def test_every_declared_symbol_is_exported(lib, header, least):
    syms = header_symbols(header)
    assert len(syms) >= least
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(BIN, lib)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_reference_abi_is_complete():
    """The 63 entry points of the reference's src/lib_ipm.cu:12-83."""
    from irs_amd.microlensing import _binding
    ref = _binding.IPM_SYMBOLS[:63]
    assert len(ref) == 63 and len(set(ref)) == 63
    assert set(ref) <= set(header_symbols("lib_ipm.h"))


@pytest.fixture(scope="module")
def libs():
    from irs_amd.microlensing.IPM import lib_ipm, lib_ipm_double
    from irs_amd.microlensing.IRS import lib_irs
    return {"float": lib_ipm.lib, "double": lib_ipm_double.lib, "irs": lib_irs.lib}


def test_defaults_match_reference(libs):
    # ipm.cuh:42-64 and irs.cuh:51,59
    L = libs["double"]
    o = L.init()
    assert L.get_kappa_tot(o) == pytest.approx(0.3)
    assert L.get_kappa_star(o) == pytest.approx(0.27)
    assert L.get_mass_function(o) == b"equal"
    assert L.get_m_lower(o) == pytest.approx(0.01) and L.get_m_upper(o) == pytest.approx(50)
    assert L.get_light_loss(o) == pytest.approx(0.001)
    assert L.get_safety_scale(o) == pytest.approx(1.37)
    assert (L.get_half_length_y1(o), L.get_half_length_y2(o)) == (5.0, 5.0)
    assert (L.get_num_pixels_y1(o), L.get_num_pixels_y2(o)) == (1000, 1000)
    assert L.get_num_rays_y(o) == 1
    assert L.get_outfile_prefix(o) == b"./"
    assert L.get_precision(o) == 0
    L.ipm_free(o)
    assert libs["float"].get_precision(libs["float"].init()) == 1
    R = libs["irs"]
    r = R.init()
    assert R.get_light_loss(r) == pytest.approx(0.01) and R.get_num_rays_y(r) == 100
    R.irs_free(r)


def test_setters_and_getters_round_trip(libs):
    from irs_amd.microlensing import _binding
    for key, L in libs.items():
        o = L.init()
        for name, kind in _binding._FIELDS:
            if kind == "s":
                getattr(L, f"set_{name}")(o, b"abc.bin")
                assert getattr(L, f"get_{name}")(o) == b"abc.bin"
            elif kind == "i":
                getattr(L, f"set_{name}")(o, 7 if name != "precision" else 1)
                assert getattr(L, f"get_{name}")(o) == (7 if name != "precision" else 1)
            else:
                getattr(L, f"set_{name}")(o, 0.625)
                assert getattr(L, f"get_{name}")(o) == 0.625
        (L.irs_free if key == "irs" else L.ipm_free)(o)


def run_in_subprocess(code):
    """Run code that calls the library in a child and return its stderr (the
    C++ messages go to file descriptor 2)."""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                       env=dict(os.environ, PYTHONPATH=ROOT))
    return r.returncode, r.stdout, r.stderr


BAD_PARAMS = [
    ("set_kappa_tot(o, 0.0)", "kappa_tot must be >="),
    ("set_kappa_star(o, 0.0)", "kappa_star must be >="),
    ("set_kappa_star(o, 0.5)", "kappa_star must be <= kappa_tot"),
    ("set_theta_star(o, 0.0)", "theta_star must be >="),
    ("set_mass_function(o, b'lognormal')", "mass_function must be equal, uniform, Salpeter, Kroupa, or optical_depth"),
    ("set_m_solar(o, 0.0)", "m_solar must be >="),
    ("set_m_lower(o, 0.0)", "m_lower must be >="),
    ("set_m_upper(o, 0.001)", "m_upper must be >= m_lower"),
    ("set_light_loss(o, 0.0)", "light_loss must be >="),
    ("set_light_loss(o, 0.02)", "light_loss must be <= 0.01"),
    ("set_rectangular(o, 2)", "rectangular must be 1 (rectangular) or 0 (circular)"),
    ("set_approx(o, 2)", "approx must be 1 (approximate) or 0 (exact)"),
    ("set_safety_scale(o, 1.05)", "safety_scale must be >= 1.1"),
    ("set_half_length_y1(o, 0.0)", "half_length_y1 and half_length_y2 must both be >="),
    ("set_num_pixels_y2(o, 0)", "num_pixels_y1 and num_pixels_y2 must both be integers > 0"),
    ("set_num_rays_y(o, 0)", "num_rays_y must be an integer > 0"),
    ("set_write_stars(o, 3)", "write_stars must be 1 (true) or 0 (false)"),
    ("set_write_maps(o, 3)", "write_maps must be 1 (true) or 0 (false)"),
    ("set_write_parities(o, 3)", "write_parities must be 1 (true) or 0 (false)"),
    ("set_write_histograms(o, 3)", "write_histograms must be 1 (true) or 0 (false)"),
    ("set_expansion_order(o, 40)", "expansion_order must be 0 (automatic) or between 1 and 31"),
    ("set_precision(o, 5)", "precision must be 0 (fp64) or 1 (mixed)"),
]


@pytest.mark.parametrize("setter,message", BAD_PARAMS)
def test_validation_runs_before_the_device(setter, message):
    code = ("from irs_amd.microlensing.IPM import lib_ipm_double as m\n"
            "L = m.lib; o = L.init()\n"
            f"L.{setter}\n"
            "print(L.run(o, 0))\n")
    rc, out, err = run_in_subprocess(code)
    assert rc == 0
    assert out.strip() == "False"
    assert message in err
    assert "No HIP capable devices" not in err


def write_bin(path, n, rect, corner, theta, stars, dtype=np.float64):
    with open(path, "wb") as fh:
        np.array([n, rect], dtype=np.int32).tofile(fh)
        np.array([corner[0], corner[1], theta], dtype=dtype).tofile(fh)
        np.asarray(stars, dtype=dtype).tofile(fh)


def starfile_error(path):
    code = ("from irs_amd.microlensing.IPM import lib_ipm_double as m\n"
            "L = m.lib; o = L.init()\n"
            f"L.set_starfile(o, {str(path)!r}.encode())\n"
            "print(L.run(o, 0))\n")
    rc, out, err = run_in_subprocess(code)
    assert rc == 0 and out.strip() == "False"
    return err


def test_star_file_errors(tmp_path):
    # star.cuh:180-318
    assert "is not a .bin or .txt file" in starfile_error(tmp_path / "stars.dat")
    p = tmp_path / "zero.bin"
    write_bin(p, 0, 0, (10, 0), 1, np.zeros((0, 3)))
    assert "num_stars must be an integer > 0" in starfile_error(p)
    p = tmp_path / "rect.bin"
    write_bin(p, 1, 5, (10, 0), 1, [[0, 0, 1]])
    assert "rectangular must be 1 (rectangular) or 0 (circular)" in starfile_error(p)
    p = tmp_path / "size.bin"
    write_bin(p, 2, 0, (10, 0), 1, [[0, 0, 1]])
    assert "does not contain validly formatted" in starfile_error(p)
    p = tmp_path / "neg.bin"
    write_bin(p, 1, 1, (-1, 3), 1, [[0, 0, 1]])
    assert "corner of the star field must both be >= 0" in starfile_error(p)
    p = tmp_path / "small.bin"
    write_bin(p, 1, 0, (1.0, 0), 1, [[0, 0, 1]])
    assert "not large enough to cover the desired source plane region" in starfile_error(p)


def test_valid_star_files_pass_validation(tmp_path, gpu_available):
    if gpu_available:
        pytest.skip("checks the no-device path")
    for dtype in (np.float32, np.float64):
        p = tmp_path / f"ok_{np.dtype(dtype).name}.bin"
        write_bin(p, 1, 0, (1000.0, 0), 1, [[0, 0, 1]], dtype=dtype)
        assert "No HIP capable devices" in starfile_error(p)
    p = tmp_path / "ok.txt"
    p.write_text("0 0 1\n500 0 1\n0 -700 0.5\n")
    assert "No HIP capable devices" in starfile_error(p)


def test_staged_calls_require_order(libs):
    L = libs["double"]
    o = L.init()
    assert not L.ipm_build(o, 0)
    assert not L.ipm_shoot(o, 0, 0, -1)
    assert not L.ipm_finish(o, 0)
    L.ipm_free(o)


def test_python_mirror_value_errors():
    from irs_amd.microlensing.IPM.ipm import IPM
    m = IPM()
    for name, value in [("kappa_tot", -1), ("kappa_star", -1), ("theta_star", -1), ("m_solar", -1),
                        ("m_lower", -1), ("m_upper", -1), ("light_loss", -1), ("light_loss", 0.5),
                        ("safety_scale", 1.0), ("half_length_y1", -1), ("half_length_y2", -1),
                        ("num_pixels_y1", 0), ("num_pixels_y2", 0), ("num_rays_y", 0), ("mass_function", "x"),
                        ("expansion_order", 32), ("precision", 3), ("kernel", 2)]:
        with pytest.raises(ValueError):
            setattr(m, name, value)
    with pytest.raises(ValueError):
        IPM(kappa_tot=0.3, kappa_star=0.4)
    with pytest.raises(ValueError):
        IPM(m_lower=2.0, m_upper=1.0)
    m2 = IPM(kappa_tot=0.5, smooth_fraction=0.2)
    assert m2.kappa_star == pytest.approx(0.4)
    assert m2.smooth_fraction == pytest.approx(0.2)
    assert m2.mu_ave == pytest.approx(1 / (0.25 - 0.09))
    assert m2.pixel_scales == (0.01, 0.01)


def test_python_run_without_gpu_raises(gpu_available):
    if gpu_available:
        pytest.skip("checks the no-device path")
    from irs_amd.microlensing.IPM.ipm import IPM
    from irs_amd.microlensing.IRS.irs import IRS
    with pytest.raises(Exception, match="Error running IPM"):
        IPM(kappa_tot=0.3, shear=0.3, kappa_star=0.3, num_pixels_y1=10, num_pixels_y2=10).run()
    with pytest.raises(Exception, match="Error running IRS"):
        IRS(kappa_tot=0.3, shear=0.3, kappa_star=0.3, num_pixels_y1=10, num_pixels_y2=10).run()
