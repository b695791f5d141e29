"""Number of caustic crossings (ncc_functions.cuh, ncc.cuh).

CPU: the oracle (oracle/ncc.py) restates the reference's per-crossing loop
and its 2 x 2 reductions; it is pinned by the winding numbers of closed
clockwise curves about every (sub)pixel centre, computed by angle sums,
including curves that leave the map through each edge; evaluating the line
intersections in the reference's log-space form gives the same maps; and a
planted half-pixel shift of the crossings is caught. lib_ncc.so validates
its parameters and its caustics file before touching the GPU.

GPU: the library's maps are identical to the oracle's (integer counts, bit
for bit) for synthetic curves with axis-aligned edges, vertices on pixel-centre
rows, NaN and repeated points, for float and double caustics files, and for
caustics the critical-curve finder computes; histograms and saved files read
back; a counter-clockwise curve is rejected as the reference rejects it.
"""

import ctypes
import os

import numpy as np
import pytest

from oracle import ncc as ON


def star_curve(rng, centre, r0, n=40, clockwise=True):
    t = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = r0 * (1 + 0.4 * rng.uniform(-1, 1, n))
    z = centre + r * np.exp(1j * t)
    return z[::-1] if clockwise else z


def curve_set(seed=3):
    """Closed clockwise curves, four of them leaving the map (left, top, right, corner)."""
    rng = np.random.default_rng(seed)
    curves = [star_curve(rng, complex(*rng.uniform(-4, 4, 2)), rng.uniform(0.5, 3)) for _ in range(6)]
    curves += [star_curve(rng, complex(-5.5, 0.3), 2.0), star_curve(rng, complex(2.0, 5.2), 1.5),
               star_curve(rng, complex(6.5, -1.0), 2.5), star_curve(rng, complex(-5.0, -5.0), 1.5)]
    return curves


def as_rows(curves):
    """One caustics row per curve, the first point repeated at its end, padded by repeating the last."""
    rows = [np.concatenate([c, c[:1]]) for c in curves]
    n = max(len(r) for r in rows)
    return np.array([np.concatenate([r, np.repeat(r[-1:], n - len(r))]) for r in rows])


def pixel_centres(center, hl, npx, os_):
    s = 1 << os_
    fx, fy = npx[0] * s, npx[1] * s
    xs = center[0] - hl[0] + (np.arange(fx) + 0.5) * 2 * hl[0] / fx
    ys = center[1] + hl[1] - (np.arange(fy) + 0.5) * 2 * hl[1] / fy
    return np.meshgrid(xs, ys)


GRID = ((0.3, -0.2), (5.0, 4.0), (80, 60))


@pytest.mark.parametrize("os_", [0, 1, 2])
def test_oracle_counts_are_winding_numbers(os_):
    curves = curve_set()
    center, hl, npx = GRID
    m = ON.ncc_map(as_rows(curves), center, hl, npx, os_)
    X, Y = pixel_centres(center, hl, npx, os_)
    s = 1 << os_
    w = ON.winding_numbers(curves, X, Y).reshape(npx[1], s, npx[0], s).max(axis=(1, 3))
    assert m.max() >= 2 and np.array_equal(m, w)


def test_oracle_logspace_intersections_give_the_same_map():
    center, hl, npx = GRID
    rows = as_rows(curve_set(5))
    assert np.array_equal(ON.ncc_map(rows, center, hl, npx, 1), ON.ncc_map(rows, center, hl, npx, 1, logspace=True))


def test_oracle_counter_clockwise_curves_count_negative():
    center, hl, npx = GRID
    curves = [c[::-1] for c in curve_set()]
    fine = ON.fine_counts(as_rows(curves), center, hl, npx)
    X, Y = pixel_centres(center, hl, npx, 0)
    assert np.array_equal(fine, ON.winding_numbers(curves, X, Y)) and fine.min() < 0


def test_planted_half_pixel_shift_is_caught(monkeypatch):
    """A crossing rule that is off by half a pixel fails the winding-number check."""
    center, hl, npx = GRID
    curves = curve_set()
    X, Y = pixel_centres(center, hl, npx, 0)
    w = ON.winding_numbers(curves, X, Y)
    good = ON.to_pixel
    monkeypatch.setattr(ON, "to_pixel", lambda p, h, n: (good(p, h, n)[0] + 0.5, good(p, h, n)[1]))
    assert (ON.ncc_map(as_rows(curves), center, hl, npx, 0) != w).sum() > 0


def write_caustics(path, rows, dtype=np.float64):
    rows = np.asarray(rows)
    with open(path, "wb") as f:
        np.array([rows.shape[0], rows.shape[1]], dtype=np.int32).tofile(f)
        np.stack([rows.real, rows.imag], axis=-1).astype(dtype).tofile(f)


def lib():
    from irs_amd.microlensing.NCC import lib_ncc
    return lib_ncc.lib


def test_library_defaults_match_reference():
    """ncc.cuh:28-35."""
    L = lib()
    o = L.init()
    assert L.get_infile_prefix(o) == b"./" and L.get_outfile_prefix(o) == b"./"
    assert (L.get_center_y1(o), L.get_center_y2(o), L.get_half_length_y1(o), L.get_half_length_y2(o)) == (0, 0, 5, 5)
    assert (L.get_num_pixels_y1(o), L.get_num_pixels_y2(o), L.get_over_sample(o)) == (1000, 1000, 2)
    assert (L.get_write_maps(o), L.get_write_histograms(o)) == (1, 1)
    assert not L.get_num_crossings(o)
    L.ncc_free(o)


@pytest.mark.parametrize("setter,val,msg", [
    ("set_half_length_y1", 0.0, "half_length_y1 and half_length_y2 must both be >="),
    ("set_num_pixels_y2", 0, "num_pixels_y1 and num_pixels_y2 must both be integers > 0"),
    ("set_over_sample", -1, "over_sample must be an integer >= 0"),
    ("set_write_maps", 2, "write_maps must be 1 (true) or 0 (false)"),
    ("set_write_histograms", -1, "write_histograms must be 1 (true) or 0 (false)"),
])
def test_library_validates_before_the_device(capfd, setter, val, msg):
    L = lib()
    o = L.init()
    getattr(L, setter)(o, val)
    assert not L.run(o, 0)
    assert msg in capfd.readouterr().err
    L.ncc_free(o)


def test_caustics_file_errors(tmp_path, capfd):
    """array_functions.cuh:173-267 and ncc.cuh:194-222, all before any device call."""
    L = lib()
    o = L.init()
    L.set_infile_prefix(o, (str(tmp_path) + "/").encode())
    assert not L.run(o, 0)
    assert "Unable to read caustics from file" in capfd.readouterr().err
    path = str(tmp_path / "ccf_caustics.bin")
    with open(path, "wb") as f:
        np.array([2, 3], dtype=np.int32).tofile(f)
        np.zeros(5, dtype=np.float64).tofile(f)
    assert not L.run(o, 0)
    assert "does not contain valid single or double precision numbers" in capfd.readouterr().err
    with open(path, "wb") as f:
        np.array([0, 3], dtype=np.int32).tofile(f)
    assert not L.run(o, 0)
    assert "does not contain valid values for num_rows and num_cols" in capfd.readouterr().err
    write_caustics(path, np.zeros((4, 1), dtype=complex))
    assert not L.run(o, 0)
    assert "does not contain valid values for num_rows and num_cols" in capfd.readouterr().err
    assert not L.save(o, 0)
    L.ncc_free(o)


def test_python_mirror_value_errors():
    from irs_amd.microlensing.NCC.ncc import NCC
    for kw, msg in (({"half_length_y1": -1}, "half_length_y1 must be > 0"),
                    ({"num_pixels_y2": 0}, "num_pixels_y2 must be >= 1"),
                    ({"over_sample": -2}, "over_sample must be >= 0")):
        with pytest.raises(ValueError, match=msg):
            NCC(**kw)
    n = NCC(infile_prefix="Some/Mixed/Case/", center_y1=1.5, num_pixels_y1=7, write_maps=False)
    assert n.infile_prefix == "Some/Mixed/Case/" and n.center == (1.5, 0.0) and n.num_pixels == (7, 1000)
    assert n.pixel_scales == (10 / 7, 10 / 1000) and not n.write_maps
    n.close()


def edge_case_rows():
    """Axis-aligned rectangles (horizontal and vertical segments), one with its
    corners exactly on pixel-centre rows and columns, a row holding NaN points
    and repeated points, and a segment pair that crosses the whole grid."""
    hx, hy, nx, ny = 5.0, 4.0, 80, 60
    def on_centres(i, j):
        return complex(-hx + (i + 0.5) * 2 * hx / nx, -hy + (j + 0.5) * 2 * hy / ny)
    rect = [on_centres(10, 40), on_centres(30, 40), on_centres(30, 12), on_centres(10, 12)]   # clockwise
    box = [complex(-1, 1), complex(2, 1), complex(2, -2.5), complex(-1, -2.5)]
    long_loop = [complex(-7, 6), complex(9, 6), complex(9, -6), complex(-7, -6)]
    rows = [rect + rect[:1], box + box[:1], long_loop + long_loop[:1]]
    nan_row = [complex(0.5, 0.5), complex(np.nan, 0.2), complex(1.5, -0.5), complex(1.5, -0.5), complex(0.2, -1.0),
               complex(0.5, 0.5)]
    rows.append(nan_row)
    n = max(len(r) for r in rows)
    return np.array([r + [r[-1]] * (n - len(r)) for r in rows])


def run_ncc(tmp_path, rows, center, hl, npx, os_, dtype=np.float64, **kw):
    from irs_amd.microlensing.NCC.ncc import NCC
    write_caustics(str(tmp_path / "ccf_caustics.bin"), rows, dtype)
    n = NCC(infile_prefix=str(tmp_path) + "/", center_y1=center[0], center_y2=center[1], half_length_y1=hl[0],
            half_length_y2=hl[1], num_pixels_y1=npx[0], num_pixels_y2=npx[1], over_sample=os_,
            outfile_prefix=str(tmp_path) + "/", **kw)
    n.run()
    return n


@pytest.mark.gpu
@pytest.mark.parametrize("os_", [0, 1, 2])
def test_gpu_matches_oracle_on_curves(tmp_path, os_):
    center, hl, npx = GRID
    rows = as_rows(curve_set())
    print(f"[gpu] ncc curves over_sample={os_}", flush=True)
    n = run_ncc(tmp_path, rows, center, hl, npx, os_)
    assert n.num_caustic_crossings.shape == (npx[1], npx[0])
    assert np.array_equal(n.num_caustic_crossings, ON.ncc_map(rows, center, hl, npx, os_))
    n.close()


@pytest.mark.gpu
def test_gpu_matches_oracle_on_edge_cases(tmp_path):
    rows = edge_case_rows()
    center, hl, npx = (0.0, 0.0), (5.0, 4.0), (80, 60)
    print("[gpu] ncc edge cases", flush=True)
    for os_ in (0, 2):
        n = run_ncc(tmp_path, rows, center, hl, npx, os_)
        want = ON.ncc_map(rows, center, hl, npx, os_)
        assert want.max() >= 2 and np.array_equal(n.num_caustic_crossings, want)
        n.close()


@pytest.mark.gpu
@pytest.mark.parametrize("npx,os_", [((8200, 4), 0), ((2, 3), 9), ((3, 2), 6), ((3, 2), 7), ((2, 2), 8)])
def test_gpu_matches_oracle_on_wide_and_deep_grids(tmp_path, npx, os_):
    """Final rows wider than one workgroup's LDS row (8192 pixels) and 512 x 512
    sub-pixels per pixel take the two-pass path; 64 to 256 sub-pixels per axis
    take the one-pass path with whole-wave maxima."""
    center, hl = (0.1, 0.2), (5.0, 4.0)
    rows = as_rows(curve_set(7))
    print(f"[gpu] ncc grid {npx} over_sample={os_}", flush=True)
    n = run_ncc(tmp_path, rows, center, hl, npx, os_)
    want = ON.ncc_map(rows, center, hl, npx, os_)
    assert want.max() >= 1 and np.array_equal(n.num_caustic_crossings, want)
    n.close()


@pytest.mark.gpu
def test_gpu_reads_float_caustics(tmp_path):
    center, hl, npx = (-0.5, 0.25), (6.0, 5.0), (50, 70)
    rows = as_rows(curve_set(11))
    print("[gpu] ncc float file", flush=True)
    n = run_ncc(tmp_path, rows, center, hl, npx, 1, dtype=np.float32)
    r32 = rows.real.astype(np.float32).astype(np.float64) + 1j * rows.imag.astype(np.float32).astype(np.float64)
    assert np.array_equal(n.num_caustic_crossings, ON.ncc_map(r32, center, hl, npx, 1))
    n.close()


@pytest.mark.gpu
def test_gpu_matches_oracle_on_ccf_caustics(tmp_path):
    """The caustics the critical-curve finder writes, counted by the library and the oracle."""
    from irs_amd.microlensing.CCF.ccf import CCF
    print("[gpu] ncc on ccf caustics", flush=True)
    c = CCF(kappa_tot=0.4, shear=0.3, kappa_star=0.35, num_stars=30, num_phi=60, num_branches=2, random_seed=4,
            write_stars=False, write_critical_curves=False, outfile_prefix=str(tmp_path) + "/")
    c.run()
    c.save()
    caustics = c.caustics
    c.close()
    from irs_amd.microlensing.NCC.ncc import NCC
    center, hl, npx = (0.2, -0.1), (4.0, 4.0), (120, 100)
    n = NCC(infile_prefix=str(tmp_path) + "/", center_y1=center[0], center_y2=center[1], half_length_y1=hl[0],
            half_length_y2=hl[1], num_pixels_y1=npx[0], num_pixels_y2=npx[1], over_sample=1,
            outfile_prefix=str(tmp_path) + "/")
    n.run()
    want = ON.ncc_map(caustics, center, hl, npx, 1)
    assert want.max() >= 1 and np.array_equal(n.num_caustic_crossings, want)
    n.close()


@pytest.mark.gpu
def test_gpu_histogram_and_saved_files(tmp_path):
    from irs_amd.microlensing.Util import util
    center, hl, npx = GRID
    rows = as_rows(curve_set())
    print("[gpu] ncc saved files", flush=True)
    n = run_ncc(tmp_path, rows, center, hl, npx, 2)
    n.save()
    want = ON.ncc_map(rows, center, hl, npx, 2)
    got = util.read_array(str(tmp_path / "ncc_ncc.bin"), np.int32)
    assert np.array_equal(got, want)
    lo, counts = ON.histogram(want)
    lines = [tuple(map(int, l.split())) for l in open(tmp_path / "ncc_ncc_numpixels.txt")]
    assert lines == [(lo + k, int(v)) for k, v in enumerate(counts) if v]
    info = dict(l.split() for l in open(tmp_path / "ncc_parameter_info.txt"))
    assert info["num_pixels_y1"] == "80" and info["num_pixels_y2"] == "60" and info["over_sample"] == "2"
    assert float(info["t_ncc"]) > 0 and float(info["t_reduce"]) > 0
    n.close()


@pytest.mark.gpu
def test_gpu_counter_clockwise_curves_are_rejected(tmp_path, capfd):
    """ncc.cuh:321-326: a negative count fails the run."""
    center, hl, npx = GRID
    rows = as_rows([c[::-1] for c in curve_set()])
    print("[gpu] ncc counter-clockwise", flush=True)
    with pytest.raises(Exception, match="Error running NCC"):
        run_ncc(tmp_path, rows, center, hl, npx, 0)
    assert "Number of caustic crossings should be >= 0" in capfd.readouterr().err
