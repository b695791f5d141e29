"""File helpers (irs_amd.microlensing.Util.util) against the byte layouts of
the reference's writers: star files (star.cuh:498-534), 2D arrays
(array_functions.cuh:280-316), histograms (array_functions.cuh:318-349) and
parameter files (ipm.cuh:1141-1206), and the reference's reader semantics
(microlensing/Util/util.py:5-125)."""

import struct

import numpy as np
import pytest

from irs_amd.microlensing.Util import util


@pytest.mark.parametrize("dtype,code", [(np.float32, "f"), (np.float64, "d")])
def test_star_file_layout(tmp_path, dtype, code):
    rng = np.random.default_rng(3)
    stars = rng.normal(size=(7, 3)).astype(dtype)
    f = str(tmp_path / "s.bin")
    util.write_stars(f, 7, True, (3.5, 2.25), 1.5, stars, dtype=dtype)
    raw = open(f, "rb").read()
    # independent packing of the same record
    expect = struct.pack("ii", 7, 1) + struct.pack(code * 3, 3.5, 2.25, 1.5) + struct.pack(code * 21, *stars.ravel())
    assert raw == expect
    s, rect, corner, theta = util.read_stars(f, dtype=dtype)
    assert rect is True and theta == pytest.approx(1.5)
    assert np.array_equal(s, stars) and np.allclose(corner, (3.5, 2.25))


def test_star_file_errors(tmp_path):
    with pytest.raises(ValueError):
        util.write_stars(str(tmp_path / "s.txt"), 1, False, (1, 0), 1, np.zeros((1, 3)))
    with pytest.raises(ValueError):
        util.write_stars(str(tmp_path / "s.bin"), 1, False, (1, 0), 1, np.zeros((1, 3)), dtype=np.int32)


def test_array_layout(tmp_path):
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    f = str(tmp_path / "a.bin")
    with open(f, "wb") as fh:
        fh.write(struct.pack("ii", 3, 4) + struct.pack("f" * 12, *a.ravel()))
    assert np.array_equal(util.read_array(f, np.float32), a)
    g = str(tmp_path / "b.bin")
    util.write_array(g, a)
    assert open(g, "rb").read() == open(f, "rb").read()
    c = str(tmp_path / "c.bin")
    with open(c, "wb") as fh:
        fh.write(struct.pack("ii", 2, 2) + struct.pack("d" * 8, *range(8)))
    z = util.read_array(c, np.float64, is_complex=True)
    assert z.shape == (2, 2, 2) and z[1, 1, 1] == 7


def test_params_and_histograms(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("kappa_tot 0.45\nnum_stars 1282809\nmass_function kroupa\nlight_loss 1e-3\n")
    p = util.read_params(str(f))
    assert p == {"kappa_tot": 0.45, "num_stars": 1282809, "mass_function": "kroupa", "light_loss": 0.001}
    assert isinstance(p["num_stars"], int)
    h = tmp_path / "h.txt"
    h.write_text("1000 5\n1001 7\n")
    assert util.read_hist(str(h)).tolist() == [[1000, 5], [1001, 7]]
