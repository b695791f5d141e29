"""Every file the map makers write, checked on the GPU against what the host
computes from the arrays the library returned.

  * IPM histogram files (ipm.cuh:950-1120, 1229-1284): the linear histogram
    bins mu * 1000 and the log10 histogram log10(mu) * 1000, rounded half away
    from zero (array_functions.cuh:114-160), for the total, minima and saddles
    maps of both libraries; "value count" lines for the non-empty bins
    (array_functions.cuh:318-349). Pixels of zero magnification have no log10
    and are left out of the log10 files (the reference's kernel would index its
    histogram with -inf there).
  * IRS save() (irs.cuh:941-1222): ray-count maps, count histograms, stars and
    the parameter file.
  * The ipm and irs programs (src/main_ipm_irs.cu:169-606) write the maps the
    libraries compute for the same parameters.
"""

import os
import subprocess

import numpy as np
import pytest

from common import C1_SMALL, from_starfile, write_starfile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "irs_amd", "bin")


def say(msg):
    print(msg, flush=True)


def llround(v):
    """C's llround: to nearest, halves away from zero (exact for |v| < 2^52)."""
    v = np.asarray(v, dtype=np.float64)
    t = np.trunc(v)
    return (t + np.where(np.abs(v - t) >= 0.5, np.sign(v), 0.0)).astype(np.int64)


def expected_hist(values):
    """(value, count) rows of the non-empty bins, in increasing value."""
    u, c = np.unique(values, return_counts=True)
    return np.column_stack([u, c]).astype(np.int64)


def read_hist(path):
    return np.loadtxt(path, dtype=np.int64, ndmin=2)


@pytest.fixture(scope="module")
def stars(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("out") / "stars.bin"
    write_starfile(oracle, path, C1_SMALL, seed=4)
    return path


@pytest.mark.parametrize("is_double", [True, False])
def test_ipm_histogram_files_are_exact(tmp_path, stars, is_double):
    from irs_amd.microlensing.IPM.ipm import IPM
    prefix = str(tmp_path) + "/"
    kw = from_starfile(C1_SMALL, stars)
    kw.update(write_stars=False, write_maps=False, write_parities=True, write_histograms=True, outfile_prefix=prefix)
    say(f"[gpu] IPM histograms, is_double={is_double}")
    m = IPM(**kw, is_double=is_double, expansion_order=12)
    m.run()
    m.save()
    maps = {"": m.magnifications, "_minima": m.magnifications_minima, "_saddles": m.magnifications_saddles}
    for suffix, mp in maps.items():
        v = np.asarray(mp, dtype=np.float64).ravel()
        lin = read_hist(prefix + f"ipm_mags_numpixels{suffix}.txt")
        np.testing.assert_array_equal(lin, expected_hist(llround(v * 1000.0)), err_msg=f"linear{suffix}")
        pos = v[v > 0]
        log = read_hist(prefix + f"ipm_log_mags_numpixels{suffix}.txt")
        np.testing.assert_array_equal(log, expected_hist(llround(np.log10(pos) * 1000.0)), err_msg=f"log{suffix}")
        say(f"[hist{suffix}] {len(lin)} linear bins, {len(log)} log bins, {len(v) - len(pos)} zero pixels")
    m.close()


@pytest.mark.parametrize("is_double", [True, False])
def test_log_histograms_leave_out_zero_pixels(tmp_path, is_double):
    """One negligible star far outside the field: every source position has a
    single image, a minimum, so the saddle map is zero everywhere. Its linear
    histogram is the one bin "0 n1*n2"; its log10 histogram file holds no bin
    (a zero pixel has no log10), and the total and minima files are those of
    the returned maps."""
    from irs_amd.microlensing.IPM.ipm import IPM
    f = tmp_path / "one_star.bin"
    with open(f, "wb") as fh:
        np.array([1, 0], dtype=np.int32).tofile(fh)
        np.array([1.0e4, 0.0, 1.0], dtype=np.float64).tofile(fh)
        np.array([1.0e4, 0.0, 1.0e-12], dtype=np.float64).tofile(fh)
    prefix = str(tmp_path) + "/"
    say(f"[gpu] one-star histograms, is_double={is_double}")
    m = IPM(kappa_tot=0.3, shear=0.2, starfile=str(f), half_length_y1=5, half_length_y2=5, num_pixels_y1=40,
            num_pixels_y2=30, num_rays_y=9, expansion_order=3, is_double=is_double, write_stars=False,
            write_maps=False, write_parities=True, write_histograms=True, outfile_prefix=prefix)
    m.run()
    m.save()
    assert not m.magnifications_saddles.any() and (m.magnifications_minima > 0).all()
    np.testing.assert_array_equal(read_hist(prefix + "ipm_mags_numpixels_saddles.txt"), [[0, 40 * 30]])
    assert os.path.getsize(prefix + "ipm_log_mags_numpixels_saddles.txt") == 0
    for suffix, mp in (("", m.magnifications), ("_minima", m.magnifications_minima)):
        v = np.asarray(mp, dtype=np.float64).ravel()
        np.testing.assert_array_equal(read_hist(prefix + f"ipm_log_mags_numpixels{suffix}.txt"),
                                      expected_hist(llround(np.log10(v) * 1000.0)), err_msg=suffix)
    m.close()


def test_irs_save_reads_back(tmp_path, stars):
    from irs_amd.microlensing.IRS.irs import IRS
    from irs_amd.microlensing.Util import util
    prefix = str(tmp_path) + "/"
    kw = from_starfile(C1_SMALL, stars)
    kw.update(write_stars=True, write_maps=True, write_parities=True, write_histograms=True, outfile_prefix=prefix,
              num_rays_y=16)
    say("[gpu] IRS save")
    r = IRS(**kw, expansion_order=12)
    r.run()
    r.save()
    for suffix, mp in (("", r.num_rays), ("_minima", r.num_rays_minima), ("_saddles", r.num_rays_saddles)):
        got = util.read_array(prefix + f"irs_magnifications{suffix}.bin", np.int32)
        np.testing.assert_array_equal(got, mp)
        h = read_hist(prefix + f"irs_numrays_numpixels{suffix}.txt")
        np.testing.assert_array_equal(h, expected_hist(np.asarray(mp).ravel()))
        assert h[:, 1].sum() == mp.size
    s, rect, corner, theta = util.read_stars(prefix + "irs_stars.bin", dtype=np.float32)
    np.testing.assert_array_equal(s, r.stars)
    assert not rect and theta == 1.0
    p = util.read_params(prefix + "irs_parameter_info.txt")
    assert p["num_stars"] == r.num_stars and p["num_rays_y"] == 16
    assert p["num_rays_shot"] == r.num_rays_shot and p["num_rays_received"] == r.num_rays_received
    assert p["num_rays_received"] == int(r.num_rays.sum())
    assert p["mean_num_rays_y"] == pytest.approx(r.num_rays_received / r.num_rays.size, rel=1e-5)
    r.close()


def _run(prog, args):
    say(f"[gpu] {prog} {' '.join(args)}")
    res = subprocess.run([os.path.join(BIN, prog), *args], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return res


COMMON = ["-k", "0.3", "-y", "0.3", "-ks", "0.25", "-hly1", "5", "-hly2", "5", "-npy1", "100", "-npy2", "100",
          "-rs", "5", "-ws", "1", "-wm", "1", "-wh", "1", "-wp", "1", "-v", "0", "--expansion_order", "12"]
LIB = dict(kappa_tot=0.3, shear=0.3, kappa_star=0.25, half_length_y1=5, half_length_y2=5, num_pixels_y1=100,
           num_pixels_y2=100, random_seed=5, write_stars=False, write_maps=False, write_histograms=False,
           write_parities=True, expansion_order=12)


def test_irs_program_matches_library(tmp_path):
    """irs (the reference's only IRS entry point) writes the maps lib_irs.so
    computes: integer counts, identical."""
    from irs_amd.microlensing.IRS.irs import IRS
    from irs_amd.microlensing.Util import util
    _run("irs", COMMON + ["-nry", "25", "-o", str(tmp_path) + "/"])
    r = IRS(**LIB, num_rays_y=25)
    r.run()
    for suffix, mp in (("", r.num_rays), ("_minima", r.num_rays_minima), ("_saddles", r.num_rays_saddles)):
        np.testing.assert_array_equal(util.read_array(str(tmp_path / f"irs_magnifications{suffix}.bin"), np.int32), mp)
    s, _, _, _ = util.read_stars(str(tmp_path / "irs_stars.bin"), dtype=np.float32)
    np.testing.assert_array_equal(s, r.stars)
    p = util.read_params(str(tmp_path / "irs_parameter_info.txt"))
    assert p["num_rays_received"] == r.num_rays_received
    r.close()


def test_ipm_program_matches_library(tmp_path):
    """ipm (float values, mixed precision) writes the map lib_ipm.so computes,
    up to the order of its FP32 map additions."""
    from irs_amd.microlensing.IPM.ipm import IPM
    from irs_amd.microlensing.Util import util
    _run("ipm", COMMON + ["-nry", "4", "-o", str(tmp_path) + "/"])
    m = IPM(**LIB, num_rays_y=4)
    m.run()
    for suffix, mp in (("", m.magnifications), ("_minima", m.magnifications_minima),
                       ("_saddles", m.magnifications_saddles)):
        got = util.read_array(str(tmp_path / f"ipm_magnifications{suffix}.bin"), np.float32)
        ref = np.asarray(mp, dtype=np.float64)
        e = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * max(ref.max(), 1e-30))
        assert e.max() < 2e-5, (suffix, e.max())
    s, _, _, _ = util.read_stars(str(tmp_path / "ipm_stars.bin"), dtype=np.float32)
    np.testing.assert_array_equal(s, m.stars)
    m.close()
