"""The command-line programs irs_amd/bin/ipm, ipm_double, irs, ccf and ncc:
the reference's options, messages and exit codes (src/main_ipm_irs.cu:41-606,
src/main_ccf.cu:29-523, src/main_ncc.cu:24-298)."""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "irs_amd", "bin")


def run(prog, *args, cwd=None):
    r = subprocess.run([os.path.join(BIN, prog), *map(str, args)], capture_output=True, text=True, cwd=cwd)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("prog", ["ipm", "ipm_double", "irs", "ccf"])
def test_help_and_syntax(prog):
    rc, out, _ = run(prog, "-h")
    assert rc == 255 and "Usage:" in out and "--kappa_tot" in out and "--outfile_prefix" in out
    rc, _, err = run(prog, "-k")
    assert rc == 255 and "Error. Invalid input syntax." in err
    rc, _, err = run(prog, "-zz", "1")
    assert rc == 255 and "Unknown option -zz" in err
    rc, _, err = run(prog, "-k", "abc")
    assert rc == 255 and "Error. Invalid kappa_tot input." in err
    rc, _, err = run(prog, "-s", "1.5")
    assert rc == 255 and "smooth_fraction must be < 1" in err
    rc, _, err = run(prog, "-s", "-0.5")
    assert rc == 255 and "smooth_fraction must be >= 0" in err


def test_defaults_in_help():
    # ipm.cuh:42-64 and irs.cuh:50,58
    _, out, _ = run("ipm", "-h")
    assert "light_loss. Default value: 0.001" in out and "num_rays_y. Default value: 1" in out
    _, out, _ = run("irs", "-h")
    assert "light_loss. Default value: 0.01" in out and "num_rays_y. Default value: 100" in out


def test_option_semantics(gpu_available):
    if gpu_available:
        pytest.skip("checks the no-device path")
    rc, out, err = run("ipm", "-v", "2", "-k", "0.45", "-s", "0.2", "-mf", "Kroupa", "-npy1", "10", "-npy2", "10")
    assert "mass_function set to: kroupa" in out
    assert "kappa_star set to: 0.36" in out
    assert "No HIP capable devices" in err and rc == 255
    # -ks wins over -s; a star file wins over every star-field option
    _, out, _ = run("ipm", "-v", "2", "-s", "0.5", "-ks", "0.1")
    assert "kappa_star set to: 0.1" in out and "smooth_fraction set to" not in out
    _, out, _ = run("ipm", "-v", "2", "-sf", "stars.bin", "-ks", "0.1", "-mf", "salpeter", "-k", "0.4")
    assert "kappa_star set to" not in out and "mass_function set to" not in out and "kappa_tot set to: 0.4" in out


def test_ccf_options():
    # ccf.cuh:42-62 defaults; the CCF options of main_ccf.cu:29-55; validation before the device
    _, out, _ = run("ccf", "-h")
    for text in ("num_stars. Default value: 137", "num_phi. Default value: 100", "num_branches. Default value: 1",
                 "--write_critical_curves", "--write_length_scales"):
        assert text in out
    rc, _, err = run("ccf", "-np", "7", "-v", "0")
    assert rc == 255 and "num_phi must be an even integer > 0" in err
    rc, _, err = run("ccf", "-np", "10", "-nb", "3", "-v", "0")
    assert rc == 255 and "num_phi must be a multiple of 2 * num_branches" in err
    rc, out, _ = run("ccf", "-v", "2", "-sf", "stars.bin", "-ns", "5", "-ks", "0.1", "-np", "7")
    assert "num_stars set to" not in out and "kappa_star set to" not in out and "num_phi set to: 7" in out


def test_ncc_options(tmp_path):
    # ncc.cuh:28-35 defaults and the options of main_ncc.cu:24-39; errors before the device
    rc, out, _ = run("ncc", "-h")
    assert rc == 255
    for text in ("--infile_prefix", "--center_y1", "half_length_y2. Default value: 5", "num_pixels_y1. Default value: 1000",
                 "over_sample. Default value: 2", "--write_histograms", "--outfile_prefix"):
        assert text in out
    rc, _, err = run("ncc", "-os")
    assert rc == 255 and "Error. Invalid input syntax." in err
    rc, _, err = run("ncc", "-k", "0.3")
    assert rc == 255 and "Unknown option -k" in err
    rc, _, err = run("ncc", "-npy1", "abc")
    assert rc == 255 and "Error. Invalid num_pixels_y1 input." in err
    rc, _, err = run("ncc", "-os", "-1", "-v", "0")
    assert rc == 255 and "over_sample must be an integer >= 0" in err
    rc, out, err = run("ncc", "-v", "2", "-ip", str(tmp_path) + "/", "-cy1", "0.5")
    assert rc == 255 and "center_y1 set to: 0.5" in out and "Unable to read caustics from file" in err


@pytest.mark.gpu
def test_cli_ncc_matches_library(tmp_path):
    """ncc writes the map and histogram the library computes for the same caustics and parameters."""
    from irs_amd.microlensing.NCC.ncc import NCC
    from irs_amd.microlensing.Util import util
    t = np.linspace(0, 2 * np.pi, 61)[::-1]
    rows = np.array([1.5 * np.exp(1j * t), 0.5 + 1j + 0.8 * np.exp(1j * t)])
    with open(tmp_path / "ccf_caustics.bin", "wb") as f:
        np.array(rows.shape, dtype=np.int32).tofile(f)
        np.stack([rows.real, rows.imag], axis=-1).astype(np.float64).tofile(f)
    print("[gpu] cli ncc", flush=True)
    out_dir = tmp_path / "cli"
    out_dir.mkdir()
    rc, out, err = run("ncc", "-ip", str(tmp_path) + "/", "-hly1", "3", "-hly2", "2.5", "-npy1", "90", "-npy2", "70",
                       "-os", "1", "-o", str(out_dir) + "/", "-v", "0")
    assert rc == 0, err
    n = NCC(infile_prefix=str(tmp_path) + "/", half_length_y1=3, half_length_y2=2.5, num_pixels_y1=90,
            num_pixels_y2=70, over_sample=1, write_maps=False, write_histograms=False)
    n.run()
    got = util.read_array(str(out_dir / "ncc_ncc.bin"), np.int32)
    assert np.array_equal(got, n.num_caustic_crossings) and got.max() == 2
    assert os.path.exists(out_dir / "ncc_ncc_numpixels.txt")
    n.close()


@pytest.mark.gpu
def test_cli_ccf_matches_library(tmp_path):
    """ccf writes the critical curves and caustics the library computes for the same parameters."""
    from irs_amd.microlensing.CCF.ccf import CCF
    args = ["-k", "0.4", "-y", "0.3", "-ks", "0.35", "-ns", "12", "-np", "10", "-rs", "9", "-wls", "1",
            "-o", str(tmp_path) + "/", "-v", "0"]
    print("[gpu] cli ccf", flush=True)
    rc, out, err = run("ccf", *args)
    assert rc == 0, err
    c = CCF(kappa_tot=0.4, shear=0.3, kappa_star=0.35, num_stars=12, num_phi=10, random_seed=9,
            write_length_scales=True, write_stars=False)
    c.run()
    for name, arr in (("ccf_ccs.bin", c.critical_curves), ("ccf_caustics.bin", c.caustics)):
        vals = np.fromfile(str(tmp_path / name), dtype=np.float64, offset=8).reshape(arr.shape)
        assert np.array_equal(vals, arr)
    mu = np.fromfile(str(tmp_path / "ccf_mu_length_scales.bin"), dtype=np.float64, offset=8)
    assert np.array_equal(mu, c.mu_length_scales.ravel())
    c.close()


@pytest.mark.gpu
def test_cli_map_matches_library(tmp_path):
    """ipm writes the same map the library computes for the same parameters."""
    from irs_amd.microlensing.IPM.ipm import IPM
    from irs_amd.microlensing.Util import util
    args = ["-k", "0.3", "-y", "0.3", "-ks", "0.3", "-hly1", "5", "-hly2", "5", "-npy1", "100", "-npy2", "100",
            "-nry", "4", "-rs", "5", "-ws", "1", "-wm", "1", "-wh", "1", "-o", str(tmp_path) + "/", "-v", "0",
            "--expansion_order", "12"]
    print("[gpu] cli ipm_double", flush=True)
    rc, out, err = run("ipm_double", *args)
    assert rc == 0, err
    got = util.read_array(str(tmp_path / "ipm_magnifications.bin"), np.float64)
    stars, _, _, _ = util.read_stars(str(tmp_path / "ipm_stars.bin"), dtype=np.float64)
    m = IPM(kappa_tot=0.3, shear=0.3, kappa_star=0.3, half_length_y1=5, half_length_y2=5, num_pixels_y1=100,
            num_pixels_y2=100, num_rays_y=4, random_seed=5, write_stars=False, write_maps=False, write_histograms=False,
            is_double=True, expansion_order=12)
    m.run()
    assert np.array_equal(stars, m.stars)
    e = np.abs(got - m.magnifications) / np.maximum(np.abs(m.magnifications), 1e-12)
    assert np.max(e) < 1e-10
    m.close()
