"""GPU parity of the HIP engine against the FP64 oracle, through the C ABI.

Every check runs the engine on the same star field as the oracle (the
engine's own field, in tree order, handed to the oracle) and compares:
  * the star field with the oracle's Philox generator (same seed),
  * the quadtree node for node and list for list,
  * the FMM coefficients,
  * the lens equation at random points (tree and direct sums),
  * whole IPM maps (fast FP64, reference-faithful FP64, mixed),
  * IRS maps, with bit-exact binning of the engine's own ray positions.
Tolerances are stated at each assertion.
"""

import sys

import numpy as np
import pytest

from common import C1_SMALL, C2_SMALL, C3_FULL, Mirror, from_starfile, random_points, rel_err, write_starfile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_ipm(params, **extra):
    from irs_amd.microlensing.IPM.ipm import IPM
    kw = dict(params)
    kw.update(extra)
    say(f"[gpu] IPM run: {extra}")
    m = IPM(**kw)
    m.run()
    return m


@pytest.fixture(scope="module")
def c1_stars(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("c1") / "c1_stars.bin"
    stars = write_starfile(oracle, path, C1_SMALL, seed=1)
    return path, stars


@pytest.fixture(scope="module")
def c1_fp64(c1_stars):
    return run_ipm(from_starfile(C1_SMALL, c1_stars[0]), is_double=True, expansion_order=16)


@pytest.fixture(scope="module")
def c1_mirror(oracle, c1_fp64):
    return Mirror(oracle, c1_fp64)


def test_star_field_matches_philox_oracle(oracle):
    m = run_ipm(C1_SMALL, is_double=True, expansion_order=16)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                              (100, 100), 4)
    assert m.num_stars == d.num_stars
    ref = oracle.generate_stars(d.num_stars, 0, d.corner, "equal", d.m_lower, d.m_upper, 1.0, 1)
    got = m.stars
    a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    b = got[np.lexsort((got[:, 1], got[:, 0]))]
    # same Philox stream; device and host sin/cos/sqrt may differ in the last bits
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a[:, :2]))


@pytest.mark.parametrize("mf,ml,mu,rect", [("uniform", 0.01, 50.0, 0), ("salpeter", 0.1, 1.0, 1), ("kroupa", 0.08, 1.5, 0),
                                           ("kroupa", 0.01, 50.0, 1), ("optical_depth", 0.01, 1.0, 0)])
def test_star_fields_of_every_mass_function_match_oracle(oracle, mf, ml, mu, rect):
    """Device star generation (Philox uniforms, inverse-CDF masses, disk or
    rectangle) against the oracle's for each mass function of mass_functions.cuh."""
    params = dict(C1_SMALL, mass_function=mf, m_lower=ml, m_upper=mu, rectangular=bool(rect), random_seed=5,
                  num_pixels_y1=40, num_pixels_y2=40)
    say(f"[gpu] stars {mf} {ml}-{mu} rectangular={rect}")
    m = run_ipm(params, is_double=True, expansion_order=8)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, mf, 1.0, ml, mu, 0.001, rect, 1, 1.37, (0, 0), (10, 10), (40, 40), 4)
    assert m.num_stars == d.num_stars
    ref = oracle.generate_stars(d.num_stars, rect, d.corner, mf, d.m_lower, d.m_upper, 1.0, 5)
    got = m.stars
    a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    b = got[np.lexsort((got[:, 1], got[:, 0]))]
    # same Philox stream; device and host sin/cos/sqrt/pow/log may differ in the last bits
    assert np.max(np.abs(a[:, :2] - b[:, :2])) < 1e-12 * np.max(np.abs(a[:, :2]))
    assert np.max(np.abs(a[:, 2] - b[:, 2]) / a[:, 2]) < 1e-12
    m.close()


def test_tree_matches_oracle(c1_fp64, c1_mirror):
    g_geom, g_info, g_lists = c1_fp64.export_tree()
    o_geom, o_info, o_lists = c1_mirror.tree.export()
    assert c1_fp64.tree_levels == c1_mirror.tree.levels
    assert g_geom.shape == o_geom.shape
    np.testing.assert_array_equal(g_geom, o_geom)   # same arithmetic for centres and half-lengths
    np.testing.assert_array_equal(g_info, o_info)
    np.testing.assert_array_equal(g_lists, o_lists)
    # stars came back in the oracle's tree order
    np.testing.assert_array_equal(c1_fp64.stars, c1_mirror.tree.stars())


def test_fmm_coefficients_match_oracle(c1_fp64, c1_mirror):
    for which in ("multipole", "local"):
        g = c1_fp64.export_coeffs("local" if which == "local" else "multipole")
        o = c1_mirror.tree.coeffs("local" if which == "local" else "multipole")
        scale = np.maximum(np.max(np.abs(o), axis=1, keepdims=True), 1e-300)
        err = np.max(np.abs(g - o) / scale)
        assert err < 1e-11, (which, err)


@pytest.mark.parametrize("mode,tol", [(3, 1e-11), (2, 2e-5), (1, 2e-5)])
def test_m2l_on_mfma_matches_oracle(oracle, c1_stars, c1_fp64, c1_mirror, mode, tol):
    """M2L as batched GEMMs on the matrix cores against the oracle's FP64 M2L
    (fmm.cuh:281-374): FP64 MFMA agrees to rounding; FP32 MFMA to single
    precision of the products. Mode 1 at p = 16 runs in FP32."""
    say(f"[gpu] m2l_mode {mode}")
    m = run_ipm(from_starfile(C1_SMALL, c1_stars[0]), is_double=True, expansion_order=16, m2l_mode=mode)
    g = m.export_coeffs("local")
    o = c1_mirror.tree.coeffs("local")
    scale = np.maximum(np.max(np.abs(o), axis=1, keepdims=True), 1e-300)
    err = np.max(np.abs(g - o) / scale)
    say(f"[gpu] m2l_mode {mode}: max local-coefficient error {err:.2e}")
    assert err < tol
    assert m.m2l_fp32 == (mode in (1, 2))
    # every M2L pair of the tree found its offset matrix
    _, info, _ = c1_fp64.export_tree()
    assert m.m2l_pairs == int(info[info[:, 0] >= 2, 7].sum()) > 0
    assert m.t_m2l > 0
    # the deflection stays within the reference's FMM budget of the direct sum
    z = random_points(m, 20000, seed=5)
    w = m.deflect(z)
    w_direct = oracle.deflect_direct(c1_mirror.stars, c1_mirror.lens, z)
    assert np.max(np.abs(w - w_direct)) < m.alpha_error
    m.close()


@pytest.fixture(scope="module")
def c3_starfile(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("c3") / "c3_stars.bin"
    write_starfile(oracle, path, C3_FULL, seed=3)
    return path


@pytest.mark.parametrize("p", [10, 20, 30])
def test_m2l_on_mfma_at_stress_size(oracle, c3_starfile, p):
    """The FMM stress workload at full size (3.3e6 stars, 11 levels, 7.3e7 M2L
    pairs per build) against the host oracle's FP64 FMM (fmm.cuh:21-464) on the
    same stars and tree: the engine's local coefficients, relative to each
    node's largest, match to FP64 rounding with the FP64 recurrences (mode 0)
    and FP64 MFMA (mode 3), and to single precision of the products with FP32
    MFMA (mode 2: median 1e-6, max 1e-3). Deflections at 5e4 points through
    the tree match the oracle's to 1e-9 (FP64 modes)."""
    from irs_amd.microlensing.IPM.ipm import IPM
    params = from_starfile(C3_FULL, c3_starfile)
    mir = None
    mfma_pairs = set()
    for mode in (0, 3, 2):
        say(f"[gpu] stress size, p {p}, m2l_mode {mode}")
        m = IPM(**params, is_double=True, precision=0, expansion_order=p, m2l_mode=mode)
        m.prepare()
        m.build()
        if mir is None:
            say(f"[cpu] stress size, oracle tree and FMM at p {p}")
            mir = Mirror(oracle, m)
            g_geom, g_info, _ = m.export_tree()
            o_geom, o_info, _ = mir.tree.export()
            np.testing.assert_array_equal(g_geom, o_geom)
            np.testing.assert_array_equal(g_info, o_info)
            del g_geom, g_info, o_geom, o_info
            ref = mir.tree.coeffs("local")
            scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
            z = random_points(m, 50000, seed=13)
            w_ref, _ = mir.tree.deflect(mir.lens, z)
        c = m.export_coeffs("local")
        ec = (np.abs(c - ref) / scale).max(axis=1)
        del c
        dw = float(np.abs(m.deflect(z) - w_ref).max())
        if mode:
            mfma_pairs.add(m.m2l_pairs)
        m.close()
        say(f"[gpu] p {p} mode {mode}: local coefficients vs oracle median {np.median(ec):.2e} max {ec.max():.2e}; "
            f"deflection max {dw:.2e}")
        if mode in (0, 3):
            assert ec.max() < 1e-10, ec.max()
            assert dw < 1e-9, dw
        else:
            assert np.median(ec) < 1e-6 and ec.max() < 1e-3, (np.median(ec), ec.max())
    # both MFMA modes batch the same source-target pairs
    assert len(mfma_pairs) == 1 and mfma_pairs.pop() > 0


def test_deflection_matches_oracle(oracle, c1_fp64, c1_mirror):
    z = random_points(c1_fp64, 20000, seed=3)
    w_gpu = c1_fp64.deflect(z)
    w_tree, _ = c1_mirror.tree.deflect(c1_mirror.lens, z)
    w_direct = oracle.deflect_direct(c1_mirror.stars, c1_mirror.lens, z)
    # same tree and expansion: FP64 rounding only
    assert np.max(np.abs(w_gpu - w_tree)) < 1e-10
    # the FMM itself is within the reference's error budget alpha_error (ipm.cuh:532-533)
    assert np.max(np.abs(w_tree - w_direct)) < c1_fp64.alpha_error


def test_magnification_matches_oracle(c1_fp64, c1_mirror):
    z = random_points(c1_fp64, 5000, seed=4)
    _, mu_gpu = c1_fp64.deflect(z, want_mu=True)
    _, mu_ref, _ = c1_mirror.tree.deflect(c1_mirror.lens, z, want_mu=True)
    assert np.median(rel_err(mu_gpu, mu_ref)) < 1e-10
    assert np.all(np.sign(mu_gpu) == np.sign(mu_ref))


def assert_fp64_map(got, ref, what, worst=1e-9, p99=1e-10, median=1e-12, flux=1e-11):
    """FP64 map against the FP64 oracle. Both evaluate the same algorithm on
    bit-identical lattice corners; the deflections differ in rounding (the
    device fuses multiply-adds in the near field and the Horner sums, and sums
    in another order), which moves the map by median ~1e-14 and at most ~6e-11
    on the small cases (1e-8 over the 2.99e8 cells of the full small case).
    A dropped, doubled or misplaced contribution moves a pixel by 1e-3 or more."""
    assert got.shape == ref.shape
    e = rel_err(got, ref, floor=1e-12)
    say(f"[{what}] rel err median {np.median(e):.2e}  p99 {np.percentile(e, 99):.2e}  "
        f"p99.9 {np.percentile(e, 99.9):.2e}  max {np.max(e):.2e}; flux {got.sum() / ref.sum() - 1:.2e}")
    assert np.median(e) < median
    assert np.percentile(e, 99) < p99
    assert np.max(e) < worst
    assert abs(got.sum() / ref.sum() - 1) < flux


def test_ipm_fp64_map_matches_oracle(c1_fp64, c1_mirror):
    assert_fp64_map(c1_fp64.magnifications, c1_mirror.ipm_map(), "fast fp64")


def test_ipm_reference_kernel_matches_oracle(c1_stars, c1_mirror):
    """The reference-faithful kernel forms each corner as the cell centre
    plus or minus the half-separation and divides by Smith's formula
    (ipm_functions.cuh:125-166, complex.cuh:267-280), as the reference does:
    its corners differ from the lattice points in the last bits, which the
    map amplifies to p99 8e-8, max 1.2e-6 (measured)."""
    m = run_ipm(from_starfile(C1_SMALL, c1_stars[0]), is_double=True, expansion_order=16, kernel=1)
    assert_fp64_map(m.magnifications, c1_mirror.ipm_map(), "reference fp64", worst=1e-5, p99=1e-6, median=1e-10,
                    flux=1e-9)
    m.close()


def test_ipm_mixed_map_matches_oracle(oracle, c1_stars):
    m = run_ipm(from_starfile(C1_SMALL, c1_stars[0]), is_double=False, expansion_order=16)
    assert m.precision == 1
    mir = Mirror(oracle, m, stars=c1_stars[1])   # the float ABI returns rounded stars; use the file's
    ref = mir.ipm_map()
    e = rel_err(m.magnifications, ref, floor=1e-12)
    say(f"[mixed C1] rel err median {np.median(e):.2e}  p99 {np.percentile(e, 99):.2e}  max {np.max(e):.2e}")
    assert np.median(e) < 1e-4, np.median(e)
    assert np.percentile(e, 99) < 1e-2, np.percentile(e, 99)
    m.close()


def test_c1_full_size_maps_match_oracle(oracle, tmp_path):
    """The small FP64 case at its full size: kappa_* = gamma = 0.3, equal masses,
    3048 stars (seed 0), 1000 x 1000 pixels over half-length 10, 4 x 4 rays per
    pixel (10930 x 27322 = 2.99e8 cells), p = 25. The oracle shoots the whole
    lattice in FP64 on the host (OpenMP, about ten seconds); the FP64 map is held
    to the whole-map bounds, the mixed map to the headline criterion."""
    params = dict(C1_SMALL, num_pixels_y1=1000, num_pixels_y2=1000, num_rays_y=16, random_seed=0)
    path = tmp_path / "c1_full.bin"
    stars = write_starfile(oracle, path, params, seed=0)
    assert len(stars) == 3048
    say("[gpu] C1 full size, fp64")
    a = run_ipm(from_starfile(params, path), is_double=True, expansion_order=25)
    assert a.cells_shot == 10930 * 27322
    say("[cpu] C1 full size, oracle map")
    ref = Mirror(oracle, a).ipm_map()
    assert_fp64_map(a.magnifications, ref, "C1 full fp64", worst=1e-7)
    say("[gpu] C1 full size, mixed")
    b = run_ipm(from_starfile(params, path), is_double=False, expansion_order=25)
    e = rel_err(b.magnifications, ref, floor=1e-12)
    say(f"[mixed C1 full] rel err median {np.median(e):.2e}  p99 {np.percentile(e, 99):.2e}  max {np.max(e):.2e}")
    assert np.median(e) < 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("rectangular,approx", [(0, 0), (1, 1), (1, 0)])
def test_smooth_matter_variants_match_oracle(oracle, rectangular, approx):
    """The exact circular sheet, and the rectangular field with the approximate
    (Taylor) and the exact sheet (alpha_smooth.cuh:62-236), with smooth matter
    kappa_tot - kappa_* = 0.1: deflections and FP64 maps against the oracle."""
    params = dict(C1_SMALL, kappa_tot=0.4, kappa_star=0.3, shear=0.2, rectangular=bool(rectangular),
                  approx=bool(approx), random_seed=3)
    m = run_ipm(params, is_double=True, expansion_order=16)
    mir = Mirror(oracle, m)
    z = random_points(m, 5000, seed=9)
    w_gpu = m.deflect(z)
    w_tree, _ = mir.tree.deflect(mir.lens, z)
    dw = np.max(np.abs(w_gpu - w_tree))
    say(f"[smooth r={rectangular} a={approx}] max deflection difference {dw:.2e}")
    assert dw < 1e-10
    assert_fp64_map(m.magnifications, mir.ipm_map(), f"smooth r={rectangular} a={approx}")
    m.close()


def test_off_centre_rectangular_map_matches_oracle(oracle, c1_stars):
    """A map centred off the origin, with unequal half-lengths and pixel counts:
    the FP64 IPM map and the IRS counts against the oracle's."""
    params = dict(from_starfile(C1_SMALL, c1_stars[0]), center_y1=1.3, center_y2=-0.8, half_length_y1=6.0,
                  half_length_y2=4.5, num_pixels_y1=80, num_pixels_y2=60)
    m = run_ipm(params, is_double=True, expansion_order=16)
    assert m.magnifications.shape == (60, 80)
    mir = Mirror(oracle, m)
    assert_fp64_map(m.magnifications, mir.ipm_map(), "off-centre fp64")
    m.close()
    r = run_irs(params, precision=0, expansion_order=16, num_rays_y=9)
    ref, got = Mirror(oracle, r, stars=c1_stars[1]).irs_map()
    assert r.num_rays_received == got
    diff = np.abs(r.num_rays.astype(np.int64) - ref).sum()
    say(f"[off-centre irs] {diff} rays differ")
    assert diff <= 2
    r.close()


def test_window_cells_cover_the_window(c1_fp64, c1_mirror):
    """The cells ipm_window_cells lists are all the cells that reach the window:
    re-shot by the oracle they give that window of the oracle's full map, and the
    engine's map agrees with it as whole maps do."""
    x0, y0, w, h = 30, 42, 25, 20
    say("[gpu] window cells")
    rec = c1_fp64.window_cells(x0, y0, w, h, margin=1.0)
    assert len(rec) > 0 and (rec[:, 2] != 0).all()
    full = c1_mirror.ipm_map()
    win = c1_mirror.ipm_window(rec, x0, y0, w, h)
    np.testing.assert_allclose(win, full[y0:y0 + h, x0:x0 + w], rtol=1e-12, atol=1e-13)
    # the per-pixel bounds of whole maps (a window this small does not average the flux to 1e-9)
    e = rel_err(c1_fp64.magnifications[y0:y0 + h, x0:x0 + w], win, floor=1e-12)
    say(f"[fp64 window] rel err median {np.median(e):.2e}  max {np.max(e):.2e}")
    assert np.median(e) < 1e-12 and np.percentile(e, 99) < 1e-10 and np.max(e) < 1e-9
    # a too small record buffer reports the full count
    import ctypes
    small = np.zeros((4, 3), dtype=np.int64)
    n = c1_fp64.lib.ipm_window_cells(c1_fp64.obj, x0, y0, w, h, 1.0,
                                     small.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), 4)
    assert n == len(rec)


def test_ipm_parity_maps(oracle, c1_stars, c1_mirror):
    m = run_ipm(from_starfile(C1_SMALL, c1_stars[0]), is_double=True, expansion_order=16, write_parities=True)
    tot, mn, sd = c1_mirror.ipm_map(parities=True)
    assert_fp64_map(m.magnifications_minima + 1.0, mn + 1.0, "minima")   # +1: saddle-free pixels are 0
    assert_fp64_map(m.magnifications_saddles + 1.0, sd + 1.0, "saddles")
    assert_fp64_map(m.magnifications, tot, "minima + saddles")
    m.close()


def test_ipm_c2_shape_mixed_vs_fp64(oracle, tmp_path):
    path = tmp_path / "c2_stars.bin"
    stars = write_starfile(oracle, path, C2_SMALL, seed=7)
    a = run_ipm(from_starfile(C2_SMALL, path), is_double=True, expansion_order=20)
    b = run_ipm(from_starfile(C2_SMALL, path), is_double=False, expansion_order=20)
    mir = Mirror(oracle, a)
    ref = mir.ipm_map()
    assert_fp64_map(a.magnifications, ref, "C2-shape fp64")
    mir32 = Mirror(oracle, b, stars=stars)
    ref32 = mir32.ipm_map()
    e32 = rel_err(b.magnifications, ref32, floor=1e-12)
    say(f"[mixed C2] rel err median {np.median(e32):.2e}  p99 {np.percentile(e32, 99):.2e}  max {np.max(e32):.2e}")
    # near-critical (1 - kappa - gamma = 0.05): the mixed path keeps FP64 anchors
    assert np.median(e32) < 1e-3, np.median(e32)
    a.close()
    b.close()


def test_c2_full_size_mixed_within_1e3_of_fp64():
    """The 10000^2 map near criticality at full size: kappa 0.45, gamma 0.5,
    smooth fraction 0.2, Kroupa 0.08-1.5, half-length 25, 10 rays/px, p = 30,
    4.69e10 cells. The mixed map stays within 1e-3 median relative error of the
    FP64 map of the same stars over every pixel. The map's mean magnification
    is within 3% of mu_ave = 1 / |(1 - kappa)^2 - gamma^2| (flux conservation
    over a field of 50 x 50 Einstein radii, with 0.1% light loss by design)."""
    full = dict(C2_SMALL, half_length_y1=25.0, half_length_y2=25.0, num_pixels_y1=10000, num_pixels_y2=10000,
                num_rays_y=10, random_seed=1)
    a = run_ipm(full, is_double=True, precision=0, expansion_order=30)
    fp64 = np.array(a.magnifications, dtype=np.float64)
    n_a = a.num_stars
    a.close()
    b = run_ipm(full, is_double=True, precision=1, expansion_order=30)
    assert b.num_stars == n_a
    mixed = np.array(b.magnifications, dtype=np.float64)
    b.close()
    assert mixed.shape == (10000, 10000)
    e = rel_err(mixed, fp64, floor=1e-12)
    med = float(np.median(e))
    mu_ave = 1.0 / abs((1 - 0.45) ** 2 - 0.5 ** 2)
    mean = float(fp64.mean())
    say(f"[C2 full] {n_a} stars; mixed vs fp64 median {med:.2e}, p99 {np.percentile(e, 99):.2e}; "
        f"mean mu {mean:.4f} vs mu_ave {mu_ave:.4f}")
    assert med < 1e-3, med
    assert abs(mean / mu_ave - 1) < 0.03, (mean, mu_ave)


def run_irs(params, **extra):
    from irs_amd.microlensing.IRS.irs import IRS
    kw = dict(params)
    kw.update(extra)
    say(f"[gpu] IRS run: {extra}")
    m = IRS(**kw)
    m.run()
    return m


@pytest.mark.parametrize("precision", [0, 1])
def test_irs_binning_is_bit_exact(oracle, c1_stars, precision):
    m = run_irs(from_starfile(C1_SMALL, c1_stars[0]), precision=precision, expansion_order=16, num_rays_y=9)
    h1, h2 = m.half_length_y1, m.half_length_y2
    n1, n2 = m.num_pixels_y1, m.num_pixels_y2
    rows = np.linspace(0, m.num_ray_threads[1] - 1, 12).astype(int)
    counts = np.zeros(n1 * n2, dtype=np.int64)
    for r in rows:
        w, gpu_bin = m.ray_positions(r)
        host = np.array([oracle.irs_bin(x.real - m.center_y1, x.imag - m.center_y2, h1, h2, n1, n2)
                         if np.isfinite(x) else -2 for x in w])
        np.testing.assert_array_equal(gpu_bin, host)
        ok = gpu_bin >= 0
        np.add.at(counts, gpu_bin[ok], 1)
    # the map kernel bins those rows exactly as the positions kernel does
    m.shoot(int(rows[0]), int(rows[0]) + 1)
    m.finish()
    w, gb = m.ray_positions(int(rows[0]))
    single = np.zeros(n1 * n2, dtype=np.int64)
    np.add.at(single, gb[gb >= 0], 1)
    np.testing.assert_array_equal(m.num_rays.ravel(), single)
    m.close()


def test_irs_map_matches_oracle(oracle, c1_stars):
    m = run_irs(from_starfile(C1_SMALL, c1_stars[0]), precision=0, expansion_order=16, num_rays_y=9,
                light_loss=0.001)
    mir = Mirror(oracle, m, stars=c1_stars[1])
    ref, got = mir.irs_map()
    assert m.num_rays_received == got
    assert m.num_rays.sum() == got
    # FP64 positions differ from the oracle's by rounding only; a ray can only
    # change pixel when it lands within ~1e-12 of a pixel edge
    diff = np.abs(m.num_rays.astype(np.int64) - ref).sum()
    assert diff <= 2, diff
    m.close()


def test_irs_parity_maps_match_oracle(oracle, c1_stars):
    """IRS with write_parities (irs_functions.cuh:186-197): each received ray
    also counts as a minimum (mu >= 0) or a saddle (mu < 0), in FP64 against the
    oracle's rays."""
    m = run_irs(from_starfile(C1_SMALL, c1_stars[0]), precision=0, expansion_order=16, num_rays_y=9,
                light_loss=0.001, write_parities=True)
    mir = Mirror(oracle, m, stars=c1_stars[1])
    ref, ref_min, ref_sad, got = mir.irs_map(parities=True)
    assert m.num_rays_received == got
    assert (m.num_rays_minima + m.num_rays_saddles == m.num_rays).all()
    assert m.num_rays_minima.sum() > 0 and m.num_rays_saddles.sum() > 0
    # as for the total map, a ray changes pixel only within ~1e-12 of a pixel edge; its parity
    # changes only where |1/mu| is at the rounding level, on a critical curve
    for what, g, r in (("total", m.num_rays, ref), ("minima", m.num_rays_minima, ref_min),
                       ("saddles", m.num_rays_saddles, ref_sad)):
        diff = np.abs(g.astype(np.int64) - r).sum()
        say(f"[irs parities] {what}: {diff} rays differ")
        assert diff <= 2, (what, diff)
    m.close()


def test_irs_c2_full_size_mixed_vs_fp64():
    """IRS on the 10000^2 configuration at full size (kappa 0.45, gamma 0.5,
    smooth fraction 0.2, Kroupa 0.08-1.5, 10 rays per pixel, p = 30, light loss
    0.01). The mixed and FP64 ray counts of the same stars differ only where a
    ray lands within the FP32 deflection error of a pixel boundary: their L1
    distance is below 1e-3 of all rays received. Flux conservation: the mean
    magnification is within 3% of mu_ave = 1 / |(1 - kappa)^2 - gamma^2|."""
    full = dict(C2_SMALL, half_length_y1=25.0, half_length_y2=25.0, num_pixels_y1=10000, num_pixels_y2=10000,
                num_rays_y=10, random_seed=1, light_loss=0.01)
    a = run_irs(full, precision=0, expansion_order=30)
    fp64 = np.array(a.num_rays, dtype=np.int64)
    mu64 = float(a.magnifications.mean())
    shot = a.num_rays_shot
    a.close()
    b = run_irs(full, precision=1, expansion_order=30)
    assert b.num_rays_shot == shot
    mixed = np.array(b.num_rays, dtype=np.int64)
    b.close()
    assert mixed.shape == (10000, 10000)
    l1 = float(np.abs(mixed - fp64).sum()) / float(fp64.sum())
    mu_ave = 1.0 / abs((1 - 0.45) ** 2 - 0.5 ** 2)
    say(f"[IRS C2 full] {shot} rays shot, {fp64.sum()} received; mixed vs fp64 L1 {l1:.2e}; "
        f"mean mu {mu64:.4f} vs mu_ave {mu_ave:.4f}")
    assert l1 < 1e-3, l1
    assert abs(mu64 / mu_ave - 1) < 0.03, (mu64, mu_ave)


def test_no_star_field_is_uniform(tmp_path):
    """One negligible star far outside the field: the lens map is linear, so
    every interior pixel has mu = 1 / ((1 - kappa)^2 - gamma^2)."""
    from irs_amd.microlensing.IPM.ipm import IPM
    f = tmp_path / "one_star.bin"
    with open(f, "wb") as fh:
        np.array([1, 0], dtype=np.int32).tofile(fh)
        np.array([1.0e4, 0.0, 1.0], dtype=np.float64).tofile(fh)   # corner (R, 0), theta
        np.array([1.0e4, 0.0, 1.0e-12], dtype=np.float64).tofile(fh)
    for is_double in (True, False):
        say(f"[gpu] no-star KAT is_double={is_double}")
        m = IPM(kappa_tot=0.3, shear=0.2, starfile=str(f), half_length_y1=5, half_length_y2=5, num_pixels_y1=50,
                num_pixels_y2=50, num_rays_y=9, expansion_order=3, is_double=is_double, write_stars=False,
                write_maps=False, write_histograms=False)
        m.run()
        mu = 1 / ((1 - 0.3) ** 2 - 0.2 ** 2)
        inner = m.magnifications[2:-2, 2:-2]
        tol = 1e-9 if is_double else 2e-5
        assert np.max(np.abs(inner / mu - 1)) < tol, np.max(np.abs(inner / mu - 1))
        m.close()


@pytest.mark.parametrize("is_double", [True, False])
def test_saved_files_read_back(tmp_path, c1_stars, is_double):
    """save() writes the reference's files (ipm.cuh:1122-1322): every one reads
    back through the Util helpers to exactly what the library returned."""
    from irs_amd.microlensing.IPM.ipm import IPM
    from irs_amd.microlensing.Util import util
    prefix = str(tmp_path) + "/"
    kw = from_starfile(C1_SMALL, c1_stars[0])
    kw.update(write_stars=True, write_maps=True, write_parities=True, write_histograms=True, outfile_prefix=prefix)
    say(f"[gpu] save and read back, is_double={is_double}")
    m = IPM(**kw, is_double=is_double, expansion_order=12)
    m.run()
    m.save()
    dt = np.float64 if is_double else np.float32
    assert np.array_equal(util.read_array(prefix + "ipm_magnifications.bin", dt), m.magnifications)
    assert np.array_equal(util.read_array(prefix + "ipm_magnifications_minima.bin", dt), m.magnifications_minima)
    assert np.array_equal(util.read_array(prefix + "ipm_magnifications_saddles.bin", dt), m.magnifications_saddles)
    stars, rect, corner, theta = util.read_stars(prefix + "ipm_stars.bin", dtype=dt)
    assert np.array_equal(stars, m.stars) and not rect and theta == 1.0
    p = util.read_params(prefix + "ipm_parameter_info.txt")
    assert p["num_stars"] == m.num_stars
    assert p["num_pixels_y1"] == C1_SMALL["num_pixels_y1"] and p["kappa_tot"] == pytest.approx(C1_SMALL["kappa_tot"])
    npix = C1_SMALL["num_pixels_y1"] * C1_SMALL["num_pixels_y2"]
    for suffix in ("", "_minima", "_saddles"):
        h = util.read_hist(prefix + f"ipm_mags_numpixels{suffix}.txt")
        assert h[:, 1].sum() == npix
    # the linear histogram bins mu * 1000 rounded half away from zero (ipm.cuh:1003-1010)
    h = util.read_hist(prefix + "ipm_mags_numpixels.txt")
    v = m.magnifications.astype(np.float64).ravel() * 1000.0 - float(h[0, 0])
    ref = np.bincount(np.floor(v + 0.5).astype(np.int64))
    assert np.array_equal(h[:, 1], ref[ref > 0])
    m.close()


def test_batched_maps_match_single_runs():
    """Several heterogeneous maps in one graph-captured persistent launch equal
    the same instances shot one by one (FP32 map atomics: summation order
    only), and each histogram is the 128-bin log10 histogram of its map."""
    from irs_amd.microlensing.IPM.batch import HIST_BINS, IPMBatch, draw_instances
    from irs_amd.microlensing.IPM.ipm import IPM
    params = draw_instances(6, seed=3, sides=(64, 128, 192), star_range=(1.5e3, 4.0e3), num_rays_y=3)
    batch = IPMBatch()
    objs = []
    for p in params:
        m = IPM(**p)
        m.prepare()
        m.build()
        batch.add(m)
        objs.append(m)
    say(f"[gpu] batch of {len(objs)} maps, {batch.cells} cells")
    for rep in range(2):   # the first run autotunes, the second replays the graph
        h = batch.run()
        say(f"[gpu] batch run {rep}: {batch.t_shoot:.4f} s, tile_h {batch.tile_h}, blocks/CU {batch.blocks_per_cu}")
        for k, (p, m) in enumerate(zip(params, objs)):
            single = IPM(**p)
            single.run()
            ref = single.magnifications.astype(np.float64)
            got = m.magnifications.astype(np.float64)
            e = rel_err(got, ref, floor=1e-3 * max(ref.max(), 1e-30))
            assert np.max(e) < 2e-5, (k, np.max(e))
            assert h[k].sum() == p["num_pixels_y1"] * p["num_pixels_y2"]
            lv = np.where(got > 0, np.log10(np.maximum(got, 1e-300)), -np.inf)
            b = np.clip(np.floor((lv + 2.0) * (HIST_BINS / 4.0)), 0, HIST_BINS - 1).astype(int)
            ref_h = np.bincount(b.ravel(), minlength=HIST_BINS)
            # device log10f and numpy may round a value at a bin edge differently
            assert np.abs(h[k].astype(int) - ref_h).sum() <= 4, (k, np.abs(h[k].astype(int) - ref_h).sum())
            single.close()
    batch.close()
    for m in objs:
        m.close()


def test_batch_of_256_maps_at_full_size():
    """The heterogeneous batch at full size: 256 instances (draw_instances, seed
    0; sides 256 to 2048 px, 6.0e10 cells) in one graph-captured persistent
    launch. Every histogram counts every pixel of its map once; eight instances,
    spread over the batch, equal the same instances shot alone (FP32 map
    atomics: summation order only)."""
    from irs_amd.microlensing.IPM.batch import IPMBatch, draw_instances
    from irs_amd.microlensing.IPM.ipm import IPM
    params = draw_instances(256, seed=0)
    batch = IPMBatch()
    objs = []
    for p in params:
        m = IPM(**p)
        m.prepare()
        m.build()
        batch.add(m)
        objs.append(m)
    say(f"[gpu] batch of {len(objs)} maps, {batch.cells:.3e} cells")
    h = batch.run()
    say(f"[gpu] batch: {batch.t_shoot:.3f} s, tile_h {batch.tile_h}, blocks/CU {batch.blocks_per_cu}")
    for k, p in enumerate(params):
        assert h[k].sum() == p["num_pixels_y1"] * p["num_pixels_y2"], k
    for k in range(0, 256, 32):
        p = params[k]
        single = IPM(**p)
        single.run()
        ref = single.magnifications.astype(np.float64)
        got = objs[k].magnifications.astype(np.float64)
        e = rel_err(got, ref, floor=1e-3 * max(ref.max(), 1e-30))
        say(f"[gpu] instance {k}: {p['num_pixels_y1']} px, max rel err vs single {np.max(e):.2e}")
        assert np.max(e) < 2e-5, (k, np.max(e))
        single.close()
    batch.close()
    for m in objs:
        m.close()

