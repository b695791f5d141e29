"""GPU runs checked against the oracle where the drop-in defaults, the tall
tiles of the big maps and the direct sums decide the result.

  * the automatic expansion order (ipm.cuh:848-860), which every caller of the
    reference's Python binding gets (it cannot set the order);
  * the star moments of generated fields, the corner re-scaled with them and
    the tree's root half-length (star.cuh:124-161, ipm.cuh:717-754), checked by
    Mirror for every mass function;
  * forced tile heights (the 1024-row IPM and 256-row IRS tiles the 10000^2
    maps use, and odd heights) against the oracle's maps;
  * the maps' error against the oracle's direct sum (alpha_star.cuh:18-41 over
    every star) on seeded lattice rows of the full small FP64 case, and the
    10000^2 case's deflections at sampled lattice corners against the direct
    sum over all 1.28e6 stars, within alpha_error (ipm.cuh:532-533);
  * a random 512^2 window of the 10000^2 map against the oracle's re-shoot.
"""

import sys

import numpy as np
import pytest

from common import C1_SMALL, C2_SMALL, Mirror, from_starfile, rel_err, write_starfile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(params, irs=False, **extra):
    from irs_amd.microlensing.IPM.ipm import IPM
    from irs_amd.microlensing.IRS.irs import IRS
    kw = dict(params)
    kw.update(extra)
    say(f"[gpu] {'IRS' if irs else 'IPM'} run: {extra}")
    m = IRS(**kw) if irs else IPM(**kw)
    m.run()
    return m


@pytest.fixture(scope="module")
def c1_stars(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("c1g") / "c1_stars.bin"
    stars = write_starfile(oracle, path, C1_SMALL, seed=1)
    return path, stars


def assert_fp64_map(got, ref, what, p99=1e-10, worst=1e-9):
    """As the FP64 map bounds of the parity tests: the same lattice corners,
    FP64 rounding differences of their deflections only."""
    e = rel_err(got, ref, floor=1e-12)
    say(f"[{what}] rel err median {np.median(e):.2e}  p99 {np.percentile(e, 99):.2e}  max {np.max(e):.2e}")
    assert np.median(e) < 1e-12
    assert np.percentile(e, 99) < p99
    assert np.max(e) < worst
    assert abs(got.sum() / ref.sum() - 1) < 1e-11


@pytest.mark.parametrize("lib", ["double", "float", "irs"])
def test_automatic_expansion_order(oracle, c1_stars, lib):
    """No expansion_order given: the order is the reference's automatic one,
    ceil(2 log2 theta + L - log2 root_hl - log2 alpha_err) (ipm.cuh:848-860),
    evaluated by the oracle on its own tree, and the map at that order matches
    the oracle's."""
    params = from_starfile(C1_SMALL, c1_stars[0])
    if lib == "irs":
        m = make(params, irs=True, precision=0, num_rays_y=9, light_loss=0.001)
    else:
        m = make(params, is_double=(lib == "double"))
    mir = Mirror(oracle, m, stars=c1_stars[1])
    d = mir.derived
    want = oracle.expansion_order(d.theta_star, mir.tree.levels, mir.root_half_length, d.alpha_error)
    say(f"[auto p {lib}] engine {m.expansion_order}, oracle {want}, levels {m.tree_levels}")
    assert 3 <= want <= 31
    assert m.tree_levels == mir.tree.levels
    assert m.expansion_order == want == mir.p
    if lib == "double":
        assert_fp64_map(m.magnifications, mir.ipm_map(), "auto p fp64")
    elif lib == "float":
        e = rel_err(m.magnifications, mir.ipm_map(), floor=1e-12)
        say(f"[auto p mixed] rel err median {np.median(e):.2e}")
        assert np.median(e) < 1e-4
    else:
        ref, got = mir.irs_map()
        assert m.num_rays_received == got
        assert np.abs(m.num_rays.astype(np.int64) - ref).sum() <= 2
    m.close()


@pytest.mark.parametrize("mf,ml,mu,rect", [("equal", 1.0, 1.0, 0), ("uniform", 0.01, 50.0, 1),
                                           ("salpeter", 0.1, 1.0, 0), ("kroupa", 0.08, 1.5, 1),
                                           ("optical_depth", 0.01, 1.0, 0)])
@pytest.mark.parametrize("is_double", [True, False])
def test_generated_field_moments_and_corner(oracle, mf, ml, mu, rect, is_double):
    """A field drawn on the device: its moments (k_star_stats against
    star.cuh:124-161 over the returned stars), the corner re-scaled with the
    realised mean mass (ipm.cuh:717-731) and the root half-length
    (ipm.cuh:746-754) equal the oracle's; Mirror asserts each getter."""
    params = dict(C1_SMALL, mass_function=mf, m_lower=ml, m_upper=mu, rectangular=bool(rect), random_seed=9,
                  num_pixels_y1=48, num_pixels_y2=48, kappa_tot=0.4, kappa_star=0.3)
    m = make(params, is_double=is_double, expansion_order=8)
    mir = Mirror(oracle, m)
    sp = oracle.star_params(np.array(m.stars, dtype=np.float64), rect, tuple(mir.derived.corner),
                            mir.derived.theta_star)
    say(f"[moments {mf} rect={rect} double={is_double}] kappa_star_actual {m.lib.get_kappa_star_actual(m.obj):.6f} "
        f"(oracle {sp['kappa_star']:.6f}), corner ({m.lib.get_corner_x1(m.obj):.4f}, {m.lib.get_corner_x2(m.obj):.4f})")
    # the realised kappa_star is the target up to the sampling of N masses
    assert abs(sp["kappa_star"] / 0.3 - 1) < 0.2
    m.close()


def test_tall_and_odd_tiles_match_oracle(oracle, c1_stars):
    """The tile heights the 10000^2 maps run with (1024 IPM lattice rows, 256
    IRS rows), and odd heights that leave a short last tile, against the oracle
    (the small case's lattice has 1366 IPM rows and 2050 IRS rows)."""
    params = from_starfile(C1_SMALL, c1_stars[0])
    ref_map = None
    for th in (0, 1024, 77):
        m = make(params, is_double=True, expansion_order=16, tile_rows=th)
        say(f"[tiles ipm fp64] tile_rows {th}: launched {m.last_tile_rows}")
        assert m.last_tile_rows == (th or m.last_tile_rows)
        if ref_map is None:
            mir = Mirror(oracle, m)
            ref_map = mir.ipm_map()
        assert_fp64_map(m.magnifications, ref_map, f"ipm fp64 tile {th}")
        m.close()
    for th in (1024, 77):
        m = make(params, is_double=False, expansion_order=16, tile_rows=th)
        e = rel_err(m.magnifications, Mirror(oracle, m, stars=c1_stars[1]).ipm_map(), floor=1e-12)
        say(f"[tiles ipm mixed] tile_rows {th}: median {np.median(e):.2e}")
        assert m.last_tile_rows == th and np.median(e) < 1e-4
        m.close()
    irs_ref = None
    for th in (0, 256, 77, 32):
        r = make(params, irs=True, precision=0, expansion_order=16, num_rays_y=9, light_loss=0.001, tile_rows=th)
        say(f"[tiles irs] tile_rows {th}: launched {r.last_tile_rows}")
        if irs_ref is None:
            irs_ref = Mirror(oracle, r, stars=c1_stars[1]).irs_map()
        ref, got = irs_ref
        assert r.num_rays_received == got
        assert np.abs(r.num_rays.astype(np.int64) - ref).sum() <= 2
        r.close()


def test_c1_full_size_rows_against_direct_sum(oracle, tmp_path):
    """The small FP64 case at full size (3048 stars, seed 0, 1000^2 px, 16 rays
    per pixel, p = 25) on three seeded blocks of 48 lattice rows, against the
    oracle's direct sum over every star for every corner (no tree, no FMM).
    The oracle's own tree + FMM map of those rows differs from its direct sum by
    median 3e-13, p99 8e-10, max 4e-5 (the FMM truncation at p = 25 moves
    near-degenerate triangles between pixels); the bounds below sit outside
    that. The whole map is shot at the 1024-row tile height too."""
    params = dict(C1_SMALL, num_pixels_y1=1000, num_pixels_y2=1000, num_rays_y=16, random_seed=0)
    path = tmp_path / "c1_full.bin"
    write_starfile(oracle, path, params, seed=0)
    a = make(from_starfile(params, path), is_double=True, expansion_order=25)
    mir = Mirror(oracle, a)
    whole = np.array(a.magnifications)
    b = make(from_starfile(params, path), is_double=True, expansion_order=25, tile_rows=1024)
    say(f"[c1 full] tiles {a.last_tile_rows} and {b.last_tile_rows}")
    assert b.last_tile_rows == 1024
    # the 1024-row tiling changes only the order of the FP64 map additions
    e = rel_err(b.magnifications, whole, floor=1e-12)
    assert np.median(e) < 1e-12 and np.max(e) < 1e-9, (np.median(e), np.max(e))
    b.close()
    c = make(from_starfile(params, path), is_double=False, expansion_order=25)
    rng = np.random.default_rng(2024)
    n2 = a.num_ray_threads[1]
    # the lattice spans |x2| < 68.3 and only |x2| < 10 / (1 - kappa - gamma) = 25
    # reaches the map: blocks from the central third of the rows (|x2| < 22.8)
    for r0 in sorted(rng.integers(n2 // 3, 2 * n2 // 3 - 48, size=3)):
        rows = np.arange(r0, r0 + 48)
        say(f"[cpu] direct-sum rows {r0}..{r0 + 48}")
        direct = mir.ipm_map(rows=rows, use_direct=True)
        a.shoot(r0, r0 + 48)
        a.finish()
        c.shoot(r0, r0 + 48)
        c.finish()
        nz = direct > 0
        e64 = rel_err(a.magnifications[nz], direct[nz], floor=1e-12)
        e32 = rel_err(c.magnifications[nz], direct[nz], floor=1e-12)
        say(f"[direct rows {r0}] {nz.sum()} px; fp64 median {np.median(e64):.2e} p99 {np.percentile(e64, 99):.2e} "
            f"max {e64.max():.2e}; mixed median {np.median(e32):.2e} max {e32.max():.2e}")
        assert np.median(e64) < 1e-10 and np.percentile(e64, 99) < 1e-6 and e64.max() < 1e-3
        assert abs(a.magnifications.sum() / direct.sum() - 1) < 1e-9
        assert np.median(e32) < 1e-4
        # next to nothing lands outside the pixels the direct sum reaches
        assert a.magnifications[~nz].sum() < 1e-9 * direct.sum()
    a.close()
    c.close()


def test_c2_full_size_against_direct_sum_and_window(oracle):
    """The 10000^2 map (1.28e6 stars, p = 30): Mirror checks the engine's
    geometry; deflections at 2e4 seeded lattice corners against the oracle's
    direct sum over all stars: FP64 within 1e-9 (4e-11 measured, far inside
    alpha_error = 2.5e-4, ipm.cuh:532-533); mixed precision within alpha_error
    plus the FP32 rounding of the near field, which exceeds alpha_error only
    within ~0.01 of a star (2 of 2e4 corners measured); and a seeded 512^2
    window of the mixed and FP64 maps matches the oracle's FP64 re-shoot of
    every cell that reaches it (median 1.8e-6 and 1.0e-12 measured)."""
    from irs_amd.microlensing.IPM.ipm import IPM
    full = dict(C2_SMALL, half_length_y1=25.0, half_length_y2=25.0, num_pixels_y1=10000, num_pixels_y2=10000,
                num_rays_y=10, random_seed=1)
    say("[gpu] C2 full: prepare and build")
    m = IPM(**full, is_double=True, precision=1, expansion_order=30)
    m.prepare()
    m.build()
    say("[cpu] C2 full: oracle tree and FMM")
    mir = Mirror(oracle, m)
    alpha = m.alpha_error
    rng = np.random.default_rng(77)
    n1, n2 = m.num_ray_threads
    I = rng.integers(0, n1 + 1, 20000)
    J = rng.integers(0, n2 + 1, 20000)
    g = mir.grid
    z = (g.center_x[0] - g.half_length_x[0] + 2 * g.ray_half_sep[0] * I) + \
        1j * (g.center_x[1] - g.half_length_x[1] + 2 * g.ray_half_sep[1] * J)
    say("[cpu] C2 full: direct sums at 2e4 corners over every star")
    w_direct = oracle.deflect_direct(mir.stars, mir.lens, z)
    # near a star the mixed path's FP32 near field rounds the leaf-relative
    # offsets (at most ~6 leaf half-lengths h) to 2^-24: an error up to about
    # 2^-24 * 6 h * sum m / |z - z_i|^2, dominated by the nearest stars
    from scipy.spatial import cKDTree
    geom = mir.tree.export()[0]
    _, leaf = mir.tree.deflect(mir.lens, z)
    h = geom[leaf, 2]
    dist, idx = cKDTree(mir.stars[:, :2]).query(np.column_stack([z.real, z.imag]), k=64)
    S = (mir.stars[idx, 2] / dist ** 2).sum(axis=1)
    for prec in (0, 1):
        m.precision = prec
        dw = np.abs(m.deflect(z) - w_direct)
        say(f"[C2 direct] precision {prec}: |dw| median {np.median(dw):.2e} max {dw.max():.2e} (alpha_error {alpha}); "
            f"{int(np.sum(dw > alpha))} above alpha_error, max dw / (h S) {np.max(dw / (h * S)):.2e}")
        if prec == 0:
            assert dw.max() < 1e-9
        else:
            assert np.median(dw) < 1e-6
            assert np.all(dw <= alpha + 2.0 ** -21 * h * S)
            assert np.mean(dw > alpha) < 1e-3
    m.precision = 1
    say("[gpu] C2 full: mixed map")
    m.shoot()
    m.finish()
    mixed = np.array(m.magnifications)
    m.precision = 0
    m.shoot()
    m.finish()
    fp64 = np.array(m.magnifications)
    y0, x0 = (int(v) for v in np.random.default_rng(0).integers(0, 10000 - 512, size=2))
    rec = m.window_cells(x0, y0, 512, 512, margin=1.0)
    say(f"[cpu] C2 window ({x0}, {y0}): oracle re-shoot of {len(rec)} records")
    win = mir.ipm_window(rec, x0, y0, 512, 512)
    e32 = rel_err(mixed[y0:y0 + 512, x0:x0 + 512], win, floor=1e-12)
    e64 = rel_err(fp64[y0:y0 + 512, x0:x0 + 512], win, floor=1e-12)
    say(f"[C2 window] mixed median {np.median(e32):.2e} max {e32.max():.2e}; fp64 median {np.median(e64):.2e} "
        f"max {e64.max():.2e}")
    assert np.median(e32) < 1e-4 and e32.max() < 1e-2
    assert np.median(e64) < 1e-11 and e64.max() < 1e-6
    m.close()


def test_batched_instances_against_oracle(oracle, tmp_path):
    """Six heterogeneous instances (kappa, gamma, smooth fraction and map size
    drawn as for the 256-map batch, 1.5e3-4e3 Kroupa stars each) shot in one
    graph launch: each instance's FP32 map against the oracle's FP64 map of the
    whole lattice with the same stars (written by the oracle into a star file
    per instance), held to the mixed-precision bound of the single maps."""
    from irs_amd.microlensing.IPM.batch import IPMBatch, draw_instances
    from irs_amd.microlensing.IPM.ipm import IPM
    base = dict(theta_star=1.0, m_solar=1.0, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
                center_y1=0.0, center_y2=0.0)
    params = [dict(base, **p) for p in draw_instances(6, seed=21, sides=(64, 128, 192), star_range=(1.5e3, 4.0e3),
                                                      num_rays_y=3)]
    batch = IPMBatch()
    objs, stars = [], []
    for k, p in enumerate(params):
        path = tmp_path / f"inst{k}.bin"
        stars.append(write_starfile(oracle, path, p, seed=p["random_seed"], margin=1.1))
        m = IPM(**from_starfile(p, path))   # mixed precision, p = 20 (the drawn parameters)
        m.prepare()
        m.build()
        batch.add(m)
        objs.append(m)
    say(f"[gpu] batch of {len(objs)} instances, {batch.cells} cells")
    batch.run()
    for k, (p, m) in enumerate(zip(params, objs)):
        ref = Mirror(oracle, m, stars=stars[k]).ipm_map()
        e = rel_err(m.magnifications, ref, floor=1e-12)
        say(f"[batch vs oracle {k}] kappa {p['kappa_tot']:.3f} gamma {p['shear']:.3f} {p['num_pixels_y1']} px: "
            f"median {np.median(e):.2e} p99 {np.percentile(e, 99):.2e}")
        assert np.median(e) < 1e-4, (k, np.median(e))
        assert abs(m.magnifications.sum() / ref.sum() - 1) < 1e-4
    batch.close()
    for m in objs:
        m.close()
