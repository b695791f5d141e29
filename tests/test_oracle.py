"""Pins the FP64 oracle (oracle/oracle.cpp) without the reference, which cannot
be built or run here (CUDA only, no nvcc; its binaries are absent). Each test
checks the oracle against something the mathematics or a publication fixes:
published known-answer vectors, numerical quadrature, closed forms, an
independent polygon clipper, brute-force direct sums and conservation laws.
"""

import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

from common import C1_SMALL, write_starfile

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------------- RNG
def _philox_kat():
    """Random123 known-answer vectors for philox4x32-10 (kat_vectors, Salmon et al. SC'11)."""
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as fh:
        vecs = json.load(fh)["vectors"]
    word = lambda xs: tuple(int(x, 16) for x in xs)
    return [(word(v["ctr"]), word(v["key"]), word(v["out"])) for v in vecs]


PHILOX_KAT = _philox_kat()


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert tuple(oracle.philox(ctr, key)) == expect


def test_star_uniforms_in_open_closed_interval(oracle):
    u = np.array([oracle.star_uniforms(7, i) for i in range(2000)])
    assert np.all(u > 0) and np.all(u <= 1)
    # uniform: KS against U(0,1) for each of the four streams
    for k in range(4):
        assert stats.kstest(u[:, k], "uniform").pvalue > 1e-3


# ------------------------------------------------------------------ mass functions
def kroupa_pdf(m, ml, mu, ms=1.0):
    # unnormalised, continuous at 0.08 and 0.5 M_sun (kroupa.cuh:22-33)
    m0, m1 = 0.08 * ms, 0.5 * ms
    if m < m0:
        return m ** -0.3
    if m < m1:
        return m0 ** (-0.3 + 1.3) * m ** -1.3
    return m0 ** (-0.3 + 1.3) * m1 ** (-1.3 + 2.3) * m ** -2.3


def pdf_for(name, ml, mu):
    if name == "kroupa":
        return lambda m: kroupa_pdf(m, ml, mu)
    a = -2.35 if name == "salpeter" else 0.0
    return lambda m: m ** a


MASS_CASES = [("uniform", 0.01, 50.0), ("salpeter", 0.1, 1.0), ("salpeter", 0.01, 50.0), ("kroupa", 0.01, 1.5),
              ("kroupa", 0.08, 1.5), ("kroupa", 0.1, 0.4), ("kroupa", 0.6, 20.0), ("kroupa", 0.01, 0.05)]


@pytest.mark.parametrize("name,ml,mu", MASS_CASES)
def test_mass_moments_against_quadrature(oracle, name, ml, mu):
    f = pdf_for(name, ml, mu)
    pts = [p for p in (0.08, 0.5) if ml < p < mu]
    norm = integrate.quad(f, ml, mu, points=pts or None, limit=200)[0]
    m1 = integrate.quad(lambda m: m * f(m), ml, mu, points=pts or None, limit=200)[0] / norm
    m2 = integrate.quad(lambda m: m * m * f(m), ml, mu, points=pts or None, limit=200)[0] / norm
    m2l = integrate.quad(lambda m: m * m * math.log(m) * f(m), ml, mu, points=pts or None, limit=200)[0] / norm
    got = oracle.mass_moments(name, ml, mu)
    np.testing.assert_allclose(got, [m1, m2, m2l], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name,ml,mu", MASS_CASES)
def test_mass_sampler_inverts_cdf(oracle, name, ml, mu):
    f = pdf_for(name, ml, mu)
    pts = [p for p in (0.08, 0.5) if ml < p < mu]
    norm = integrate.quad(f, ml, mu, points=pts or None, limit=200)[0]
    for p in (1e-6, 0.1, 0.37, 0.5, 0.77, 0.999):
        m = oracle.mass_sample(name, p, ml, mu)
        assert ml * (1 - 1e-12) <= m <= mu * (1 + 1e-12)
        cdf = integrate.quad(f, ml, m, points=[q for q in pts if q < m] or None, limit=200)[0] / norm
        assert abs(cdf - p) < 1e-9, (p, m, cdf)


def test_equal_and_optical_depth_masses(oracle):
    np.testing.assert_allclose(oracle.mass_moments("equal", 1, 1), [1, 1, 0])
    ml, mu = 0.2, 3.0
    # two masses contributing equally to the convergence: n_l m_l = n_u m_u
    fl = mu / (ml + mu)
    mean = fl * ml + (1 - fl) * mu
    np.testing.assert_allclose(oracle.mass_moments("optical_depth", ml, mu)[:2], [mean, fl * ml ** 2 + (1 - fl) * mu ** 2])
    assert oracle.mass_sample("optical_depth", fl * 0.999, ml, mu) == ml
    assert oracle.mass_sample("optical_depth", fl * 1.001, ml, mu) == mu


def test_generated_field_distribution(oracle):
    n = 20000
    R = 50.0
    s = oracle.generate_stars(n, 0, (R, 0.0), "kroupa", 0.08, 1.5, 1.0, 3)
    r = np.hypot(s[:, 0], s[:, 1])
    assert r.max() <= R
    assert stats.kstest(r ** 2 / R ** 2, "uniform").pvalue > 1e-3      # uniform in the disk
    ang = np.arctan2(s[:, 1], s[:, 0])
    assert stats.kstest((ang + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    f = pdf_for("kroupa", 0.08, 1.5)
    norm = integrate.quad(f, 0.08, 1.5, points=[0.5])[0]
    cdf = np.vectorize(lambda m: integrate.quad(f, 0.08, m, points=[0.5] if m > 0.5 else None)[0] / norm)
    ms = np.sort(s[:, 2])[::20]
    assert stats.kstest(ms, cdf).pvalue > 1e-3
    # rectangular field
    s = oracle.generate_stars(n, 1, (30.0, 10.0), "equal", 1, 1, 1.0, 3)
    assert np.all(np.abs(s[:, 0]) <= 30) and np.all(np.abs(s[:, 1]) <= 10)
    assert stats.kstest((s[:, 0] + 30) / 60, "uniform").pvalue > 1e-3


# -------------------------------------------------------------- derived params
def test_c1_derived_geometry(oracle):
    # values worked out independently for C1 (1000^2 px, h_y = 10, 16 rays/px)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 0.01, 50.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                              (1000, 1000), 16)
    assert d.num_stars == 3048
    assert tuple(d.num_ray_threads) == (10930, 27322)
    np.testing.assert_allclose(d.half_length_x, (27.325, 68.305), rtol=1e-12)
    assert d.alpha_error == pytest.approx(1e-3)
    # geometric invariants: cells tile the shooting region, and the region
    # covers the light-loss margin (ipm.cuh:460-466)
    rhs = np.array(d.ray_half_sep)
    np.testing.assert_allclose(np.array(d.half_length_x), rhs * np.array(d.num_ray_threads), rtol=1e-12)
    margin = np.sqrt(0.3 * 1.0 / (1.0 * 0.001))
    assert d.half_length_x[0] >= (10 + margin) / 1.0 and d.half_length_x[1] >= (10 + margin) / 0.4
    # one ray per cell: cell area times the ray density is 1
    assert 4 * rhs[0] * rhs[1] * d.num_rays_x == pytest.approx(1.0)
    # mu_ave of the macro model
    assert d.mu_ave == pytest.approx(1 / (0.7 ** 2 - 0.3 ** 2))


def test_c2_star_count(oracle):
    d = oracle.derived_params(0.45, 0.5, 0.36, 1.0, "kroupa", 1.0, 0.08, 1.5, 0.001, 0, 1, 1.37, (0, 0), (25, 25),
                              (10000, 10000), 10)
    assert tuple(d.num_ray_threads) == (47266, 992578)
    # kappa_star pi R^2 = N <m> pi theta^2 at the safety-scaled radius
    R = 1.37 * np.hypot(*d.half_length_x)
    assert abs(d.num_stars - 0.36 * R ** 2 / d.mean_mass) <= 1


def test_expansion_order_closed_form(oracle):
    """ipm.cuh:848-860, p = ceil(2 log2(theta) + L - log2(root_hl) - log2(alpha_err)).
    Powers of two make every term an exact integer, so each case below moves one
    term at a time by a known step: a dropped factor 2, a flipped sign or a swapped
    argument changes at least one of them."""
    eo = oracle.expansion_order
    assert eo(1.0, 6, 64.0, 2.0 ** -10) == 10          # 0 + 6 - 6 + 10
    assert eo(2.0, 6, 64.0, 2.0 ** -10) == 12          # theta x2: +2 (the factor 2 on log2 theta)
    assert eo(0.5, 6, 64.0, 2.0 ** -10) == 8           # theta /2: -2
    assert eo(1.0, 9, 64.0, 2.0 ** -10) == 13          # three more levels: +3
    assert eo(1.0, 6, 256.0, 2.0 ** -10) == 8          # root 4x larger: -2
    assert eo(1.0, 6, 64.0, 2.0 ** -13) == 13          # error 8x smaller: +3
    assert eo(1.0, 6, 64.0, 2.0 ** -10 * 1.01) == 10   # ceil(9.9856...) = 10
    assert eo(1.0, 6, 64.0 * 1.01, 2.0 ** -10) == 10   # ceil(9.9856...) = 10
    assert eo(1.0, 6, 64.0, 2.0 ** -10 / 1.01) == 11   # ceil(10.0144...) = 11
    # the workloads, worked out by hand with log2(x) = ln(x) / ln(2):
    # small FP64 case, root_hl = 1.1 * 100.8 = 110.9: 6 - 6.7931 + 9.9658 = 9.17 -> 10
    assert eo(1.0, 6, 110.9, 1e-3) == 10
    # 10000^2 case: field radius 1076.26, root_hl = 1.1 * 1076.26 = 1183.9, alpha = 25 / 1e5:
    # 10 - 10.209 + 11.966 = 11.76 -> 12
    assert eo(1.0, 10, 1183.9, 2.5e-4) == 12
    # FMM stress case: radius 959.31, root_hl = 1055.2, alpha = 50 / (10 * 20000) = 2.5e-4, 11 levels:
    # 11 - 10.043 + 11.966 = 12.92 -> 13
    assert eo(1.0, 11, 1055.2, 2.5e-4) == 13


def test_star_params_closed_form(oracle):
    """star.cuh:124-161 over a fixture list whose moments are exact by hand."""
    stars = np.array([[0.3, -1.0, 1.0], [2.0, 0.5, 2.0], [-4.0, 1.5, 4.0]])
    ln2 = math.log(2.0)
    # circle of radius |(3, 4)| = 5, theta = 2: kappa = m_tot theta^2 / R^2 = 7 * 4 / 25
    sp = oracle.star_params(stars, 0, (3.0, 4.0), 2.0)
    assert sp["kappa_star"] == pytest.approx(28 / 25, rel=1e-15)
    assert (sp["m_lower"], sp["m_upper"]) == (1.0, 4.0)
    assert sp["mean_mass"] == pytest.approx(7 / 3, rel=1e-15)
    assert sp["mean_mass2"] == pytest.approx(21 / 3, rel=1e-15)
    # (1 ln 1 + 4 ln 2 + 16 ln 4) / 3 = 36 ln 2 / 3
    assert sp["mean_mass2_ln_mass"] == pytest.approx(12 * ln2, rel=1e-15)
    # rectangle with corner (2, 3), theta = 1: kappa = m_tot pi theta^2 / (4 c1 c2) = 7 pi / 24
    sp = oracle.star_params(stars, 1, (2.0, 3.0), 1.0)
    assert sp["kappa_star"] == pytest.approx(7 * math.pi / 24, rel=1e-15)
    # one mass only: m_lower = m_upper and <m^2 ln m> = m^2 ln m
    sp = oracle.star_params(np.tile([[1.0, 1.0, 0.5]], (10, 1)), 0, (1.0, 0.0), 1.0)
    assert (sp["m_lower"], sp["m_upper"]) == (0.5, 0.5)
    assert sp["mean_mass2_ln_mass"] == pytest.approx(-0.25 * ln2, rel=1e-15)
    assert sp["kappa_star"] == pytest.approx(5.0, rel=1e-15)


def test_rescaled_corner_closed_form(oracle):
    """ipm.cuh:717-731 and the tree root ipm.cuh:746-754."""
    # circle: direction (0.6, 0.8) kept, radius sqrt(theta^2 N <m> / kappa) = sqrt(100 * 0.25 / 0.5)
    c, root = oracle.rescaled_corner(0, (3.0, 4.0), 1.0, 100, 0.25, 0.5)
    np.testing.assert_allclose(c, (0.6 * math.sqrt(50), 0.8 * math.sqrt(50)), rtol=1e-15)
    assert root == pytest.approx(1.1 * math.sqrt(50), rel=1e-15)
    # rectangle (2, 8): shape (sqrt(1/4), sqrt(4)) times sqrt(pi theta^2 N <m> / (4 kappa)) with theta = 2
    c, root = oracle.rescaled_corner(1, (2.0, 8.0), 2.0, 100, 0.25, 0.5)
    f = math.sqrt(math.pi * 4 * 100 * 0.25 / 2.0)
    np.testing.assert_allclose(c, (0.5 * f, 2.0 * f), rtol=1e-15)
    assert root == pytest.approx(1.1 * 2.0 * f, rel=1e-15)
    # with the model's own mean mass the derived corner comes back unchanged
    for rect in (0, 1):
        d = oracle.derived_params(0.45, 0.5, 0.36, 1.0, "kroupa", 1.0, 0.08, 1.5, 0.001, rect, 1, 1.37, (0, 0),
                                  (25, 25), (10000, 10000), 10)
        c, _ = oracle.rescaled_corner(rect, tuple(d.corner), 1.0, d.num_stars, d.mean_mass, 0.36)
        np.testing.assert_allclose(c, tuple(d.corner), rtol=1e-14)


def test_star_file_branch_of_derived_params(oracle):
    """ipm.cuh:405-432 and 507-529: a field read from a file sets the star count,
    theta and the masses; a circular field is turned to the shooting region's
    corner with the file's radius, and too small a field is refused."""
    gen = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                                (1000, 1000), 16)
    n = gen.num_stars
    R = math.hypot(*gen.corner)
    # unit masses in a disk of the generated radius: the file's kappa_star is
    # N / R^2 = 0.3 by construction, so the geometry must equal the generated one
    rng = np.random.default_rng(0)
    stars = np.column_stack([rng.uniform(-R / 2, R / 2, n), rng.uniform(-R / 2, R / 2, n), np.ones(n)])
    d, ok = oracle.derived_params_file(stars, 0, (R, 0.0), 1.0, 0.3, 0.3, 0.001, 1, 1.37, (0, 0), (10, 10),
                                       (1000, 1000), 16)
    assert ok and d.num_stars == n and d.rectangular == 0
    assert d.kappa_star == pytest.approx(0.3, rel=1e-12)
    np.testing.assert_allclose(d.half_length_x, gen.half_length_x, rtol=1e-12)
    assert tuple(d.num_ray_threads) == tuple(gen.num_ray_threads)
    np.testing.assert_allclose(d.corner, gen.corner, rtol=1e-12)   # (R, 0) turned to the region's corner
    assert d.taylor_smooth == gen.taylor_smooth
    # theta and masses come from the file: theta 2 with masses 1/4 keeps kappa
    s2 = stars.copy()
    s2[:, 2] = 0.25
    d2, ok = oracle.derived_params_file(s2, 0, (R, 0.0), 2.0, 0.3, 0.3, 0.001, 1, 1.37, (0, 0), (10, 10),
                                        (1000, 1000), 16)
    assert ok and d2.theta_star == 2.0 and d2.mean_mass == 0.25 and d2.kappa_star == pytest.approx(0.3, rel=1e-12)
    # a field 10% below the safety-scaled radius is refused
    need = 1.37 * math.hypot(gen.half_length_x[0], gen.half_length_x[1])
    _, ok = oracle.derived_params_file(stars, 0, (0.9 * need, 0.0), 1.0, 0.3, 0.3, 0.001, 1, 1.37, (0, 0), (10, 10),
                                       (1000, 1000), 16)
    assert not ok
    # a rectangle keeps its corner (its kappa_star, and so the region, follow
    # from the corner) and needs each side to cover
    c = (3.0 * gen.half_length_x[0], 3.0 * gen.half_length_x[1])
    d3, ok = oracle.derived_params_file(stars, 1, c, 1.0, 0.3, 0.3, 0.001, 1, 1.37, (0, 0), (10, 10), (1000, 1000),
                                        16)
    assert ok and tuple(d3.corner) == c and d3.kappa_star == pytest.approx(n * math.pi / (4 * c[0] * c[1]))
    assert 1.37 * d3.half_length_x[0] <= c[0] and 1.37 * d3.half_length_x[1] <= c[1]
    _, ok = oracle.derived_params_file(stars, 1, (c[0], c[1] / 6), 1.0, 0.3, 0.3, 0.001, 1, 1.37, (0, 0), (10, 10),
                                       (1000, 1000), 16)
    assert not ok


# --------------------------------------------------------------------- deflection
def lens_for(oracle, kappa, gamma, ks, R, approx=1):
    return oracle.make_lens(kappa, gamma, 1.0, ks, (R, 0.0), 0, approx, 1)


def test_point_lens_closed_form(oracle):
    stars = np.array([[0.0, 0.0, 2.0]])
    lens = lens_for(oracle, 0.1, 0.2, 0.0, 10.0)
    rng = np.random.default_rng(1)
    z = rng.uniform(-3, 3, 500) + 1j * rng.uniform(-3, 3, 500)
    expect = 0.9 * z + 0.2 * np.conj(z) - 2.0 * z / np.abs(z) ** 2
    np.testing.assert_allclose(oracle.deflect_direct(stars, lens, z), expect, rtol=1e-13, atol=1e-13)
    T = oracle.Tree(stars, 11.0)
    T.fmm(5)
    w, mu, _ = T.deflect(lens, z, want_mu=True)
    np.testing.assert_allclose(w, expect, rtol=1e-13, atol=1e-13)
    # local magnification: w = (1-k) z + g zbar - m / zbar, so dw/dz = 1-k and
    # dw/dzbar = g + m / zbar^2, mu = 1 / (|dw/dz|^2 - |dw/dzbar|^2)
    det = 0.9 ** 2 - np.abs(0.2 + 2.0 / np.conj(z) ** 2) ** 2
    np.testing.assert_allclose(mu, 1 / det, rtol=1e-10)


def test_smooth_matter_circle_exact_and_approx(oracle):
    lens_a = lens_for(oracle, 0.3, 0.0, 0.2, 5.0, approx=1)
    lens_e = lens_for(oracle, 0.3, 0.0, 0.2, 5.0, approx=0)
    nostar = np.zeros((0, 3))
    z = np.array([1 + 1j, 3 - 2j, 10 + 0j, -6 + 8j])
    wa = oracle.deflect_direct(nostar, lens_a, z)
    we = oracle.deflect_direct(nostar, lens_e, z)
    np.testing.assert_allclose(wa, 0.7 * z + 0.2 * z, rtol=1e-14)
    inside = np.abs(z) <= 5
    np.testing.assert_allclose(we[inside], (0.7 * z + 0.2 * z)[inside], rtol=1e-14)
    np.testing.assert_allclose(we[~inside], (0.7 * z + 0.2 * 25 / np.conj(z))[~inside], rtol=1e-14)


def test_rectangular_smooth_approx_matches_exact_inside(oracle):
    # the Taylor form converges to the exact rectangular sheet inside the field
    c = (20.0, 12.0)
    nostar = np.zeros((0, 3))
    ex = oracle.make_lens(0.3, 0.1, 1.0, 0.2, c, 1, 0, 1)
    ap = oracle.make_lens(0.3, 0.1, 1.0, 0.2, c, 1, 1, 61)
    z = np.array([1 + 1j, 5 - 3j, -7 + 2j, 3 + 6j])
    np.testing.assert_allclose(oracle.deflect_direct(nostar, ap, z), oracle.deflect_direct(nostar, ex, z), atol=1e-6)


@pytest.fixture(scope="module")
def c1_tree(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("o") / "c1.bin"
    stars = write_starfile(oracle, path, C1_SMALL, seed=1)
    R = np.sqrt(len(stars) * stars[:, 2].mean() / 0.3)
    T = oracle.Tree(stars, 1.1 * R)
    return T, stars, R


def test_tree_invariants(c1_tree):
    T, stars, R = c1_tree
    geom, info, lists = T.export()
    leaf = info[:, 3] == 0
    # every leaf sees at most 32 stars directly (tree_node.cuh:14, 337)
    near = np.array([info[i, 5] + info[lists[i, :8][lists[i, :8] >= 0], 5].sum() for i in np.where(leaf)[0]])
    assert near.max() <= 32
    # children partition the parent's stars
    for i in np.where(~leaf)[0]:
        c0 = info[i, 2]
        assert info[c0:c0 + 4, 5].sum() == info[i, 5]
        assert info[c0, 4] == info[i, 4]
    # every star lies in the box of each node whose range holds it
    s = T.stars()
    for i in range(len(info)):
        b, n = info[i, 4], info[i, 5]
        if n:
            pts = s[b:b + n, :2]
            assert np.all(np.abs(pts - geom[i, :2]) <= geom[i, 2] * (1 + 1e-12))


def test_interaction_lists_cover_every_star_once(c1_tree):
    """For any leaf, near-field stars (own + neighbours), the M2L and P2L lists
    of the leaf and of all its ancestors partition the whole star field."""
    T, stars, R = c1_tree
    geom, info, lists = T.export()
    n_total = len(stars)
    rng = np.random.default_rng(5)
    leaves = np.where(info[:, 3] == 0)[0]
    for leaf in rng.choice(leaves, 40, replace=False):
        covered = np.zeros(n_total, dtype=np.int64)

        def add(node):
            b, n = info[node, 4], info[node, 5]
            covered[b:b + n] += 1

        add(leaf)
        for nb in lists[leaf, :8][lists[leaf, :8] >= 0]:
            add(nb)
        node = leaf
        while node >= 0:
            for f in lists[node, 8:40][lists[node, 8:40] >= 0]:
                add(f)
            node = info[node, 1]
        assert np.all(covered == 1)


@pytest.mark.parametrize("p", [10, 16, 25])
def test_fmm_matches_direct_sum_within_alpha_error(oracle, c1_tree, p):
    T, stars, R = c1_tree
    T.fmm(p)
    lens = lens_for(oracle, 0.3, 0.3, 0.3, R)
    rng = np.random.default_rng(p)
    z = rng.uniform(-27, 27, 3000) + 1j * rng.uniform(-68, 68, 3000)
    w, _ = T.deflect(lens, z)
    wd = oracle.deflect_direct(T.stars(), lens, z)
    err = np.abs(w - wd)
    # the reference sizes p so that the deflection error is below alpha_error = 1e-3 for C1 (ipm.cuh:532-533,848)
    assert err.max() < 1e-3
    # and the truncation error falls geometrically with p
    assert err.max() < {10: 5e-4, 16: 1e-6, 25: 1e-8}[p]


def test_fmm_translations_are_exact_at_high_order(oracle):
    """With few stars and p = 31 the FMM error is the truncation of a rapidly
    converging series; M2M, L2L and M2L must not add error beyond it."""
    rng = np.random.default_rng(11)
    n = 600
    r = 40 * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    stars = np.column_stack([r * np.cos(a), r * np.sin(a), rng.uniform(0.1, 2, n)])
    T = oracle.Tree(stars, 44.0)
    assert T.levels >= 3
    T.fmm(31)
    lens = lens_for(oracle, 0.2, 0.1, 0.0, 40)
    z = rng.uniform(-25, 25, 2000) + 1j * rng.uniform(-25, 25, 2000)
    w, _ = T.deflect(lens, z)
    wd = oracle.deflect_direct(T.stars(), lens, z)
    assert np.max(np.abs(w - wd)) < 1e-9


# ------------------------------------------------------------- triangle to pixels
def clip_halfplane(poly, a, b, c):
    """Keep the part of a polygon with a*x + b*y <= c (Sutherland-Hodgman, 1974)."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        fp, fq = a * p[0] + b * p[1] - c, a * q[0] + b * q[1] - c
        if fp <= 0:
            out.append(p)
        if (fp < 0 < fq) or (fq < 0 < fp):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def shoelace(poly):
    if len(poly) < 3:
        return 0.0
    x = np.array([p[0] for p in poly])
    y = np.array([p[1] for p in poly])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def pixel_areas_independent(tri, n1, n2):
    xs = [p[0] for p in tri]
    ys = [p[1] for p in tri]
    out = {}
    for X in range(max(0, int(np.floor(min(xs)))), min(n1 - 1, int(np.floor(max(xs)))) + 1):
        for Y in range(max(0, int(np.floor(min(ys)))), min(n2 - 1, int(np.floor(max(ys)))) + 1):
            poly = list(tri)
            for a, b, c in ((-1, 0, -X), (1, 0, X + 1), (0, -1, -Y), (0, 1, Y + 1)):
                poly = clip_halfplane(poly, a, b, c)
            ar = shoelace(poly)
            if ar > 0:
                out[n1 * Y + X] = ar
    return out


def test_triangle_allocation_against_independent_clipper(oracle):
    rng = np.random.default_rng(2)
    n1, n2 = 12, 9
    for k in range(300):
        c = rng.uniform(-1, 13, 2)
        tri = [tuple(c + rng.normal(0, rng.choice([0.2, 1.0, 3.0]), 2)) for _ in range(3)]
        whole = shoelace(tri)
        if whole < 1e-9:
            continue
        factor = 0.37
        idx, val = oracle.triangle_alloc(np.array(tri).ravel(), factor, n1, n2)
        got = {}
        for i, v in zip(idx, val):
            got[int(i)] = got.get(int(i), 0.0) + v
        ref = {i: factor * a / whole for i, a in pixel_areas_independent(tri, n1, n2).items()}
        keys = set(got) | set(ref)
        for i in keys:
            assert abs(got.get(i, 0.0) - ref.get(i, 0.0)) < 1e-10, (k, i, got.get(i), ref.get(i))


def test_triangle_allocation_edge_cases(oracle):
    # fully inside one pixel: the whole factor lands there
    idx, val = oracle.triangle_alloc([2.1, 3.1, 2.9, 3.2, 2.5, 3.8], 0.5, 10, 10)
    assert list(idx) == [10 * 3 + 2] and val[0] == pytest.approx(0.5)
    # outside the map: nothing
    idx, _ = oracle.triangle_alloc([-3, -3, -2, -3, -2, -2], 0.5, 10, 10)
    assert len(idx) == 0
    # straddling the map edge: only the inside fraction is kept (whole area taken before clipping)
    idx, val = oracle.triangle_alloc([-1, 0.5, 1, 0.5, 0, 0.9], 1.0, 10, 10)
    assert sum(val) == pytest.approx(0.5)
    # vertices on pixel corners and edges
    idx, val = oracle.triangle_alloc([1, 1, 3, 1, 1, 3], 2.0, 10, 10)
    assert sum(val) == pytest.approx(2.0)
    d = dict(zip(idx, val))
    assert d[10 * 1 + 1] == pytest.approx(2.0 * 1 / 2) and d[10 * 1 + 2] == pytest.approx(2.0 * 0.5 / 2)
    # degenerate (zero-area) triangle is dropped
    idx, _ = oracle.triangle_alloc([1, 1, 2, 2, 3, 3], 1.0, 10, 10)
    assert len(idx) == 0


# ------------------------------------------------------------------------ maps
def test_ipm_without_stars_is_uniform(oracle):
    """No stars: the lens map is linear and every pixel holds the macro
    magnification 1 / ((1 - kappa + kappa_star)^2 - gamma^2) (circle approx)."""
    d = oracle.derived_params(0.4, 0.25, 0.1, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (4, 4),
                              (40, 40), 9)
    T = oracle.Tree(np.zeros((0, 3)), 100.0)
    T.fmm(3)
    lens = oracle.make_lens(0.4, 0.25, 1.0, 0.1, (100.0, 0.0))
    g = oracle.make_grid(d, (0.0, 0.0), (4.0, 4.0), (40, 40))
    m = T.ipm_rows(lens, g, np.arange(d.num_ray_threads[1]))
    mu = 1 / ((1 - 0.4 + 0.1) ** 2 - 0.25 ** 2)
    np.testing.assert_allclose(m, mu, rtol=1e-12)


def test_ipm_conserves_image_area(oracle, c1_tree):
    """The map integrates the image-plane area of the shot cells: every cell
    whose four corners map inside the map deposits exactly its area."""
    T, stars, R = c1_tree
    T.fmm(16)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                              (100, 100), 4)
    lens = lens_for(oracle, 0.3, 0.3, 0.3, R)
    g = oracle.make_grid(d, (0.0, 0.0), (10.0, 10.0), (100, 100))
    rows = np.arange(0, d.num_ray_threads[1], 97)
    m = T.ipm_rows(lens, g, rows)
    # area of all cells of those rows, in pixel units, is an upper bound reached
    # when no triangle leaves the map; per triangle factor = 2 rhs1 rhs2 n1 n2 / (4 h1 h2)
    factor = 2 * d.ray_half_sep[0] * d.ray_half_sep[1] * 100 * 100 / (4 * 10 * 10)
    total = 2 * factor * len(rows) * d.num_ray_threads[0]
    assert 0 < m.sum() <= total * (1 + 1e-12)
    # the direct-sum map agrees with the tree map within the FMM error
    md = T.ipm_rows(lens, g, rows[:3], use_direct=True)
    mt = T.ipm_rows(lens, g, rows[:3])
    assert abs(md.sum() - mt.sum()) < 1e-6 * md.sum()


def test_ipm_window_of_all_cells_equals_the_map_window(oracle, c1_tree):
    """orc_ipm_window given every cell (as 63-cell tiles with full masks, the
    last tile cut at the lattice edge) reproduces that window of the row-shot map."""
    T, stars, R = c1_tree
    T.fmm(12)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                              (60, 50), 2)
    lens = lens_for(oracle, 0.3, 0.3, 0.3, R)
    g = oracle.make_grid(d, (0.0, 0.0), (10.0, 10.0), (60, 50))
    n1, n2 = d.num_ray_threads
    rows = np.arange(0, n2, 7)
    full = T.ipm_rows(lens, g, rows)
    recs = []
    for j in rows:
        for i0 in range(0, n1, 63):
            k = min(63, n1 - i0)
            recs.append((j, i0, (1 << k) - 1 if k < 63 else (1 << 63) - 1))
    recs = np.array([(a, b, np.int64(np.uint64(c).view(np.int64))) for a, b, c in recs], dtype=np.int64)
    x0, y0, w, h = 17, 9, 25, 30
    win = T.ipm_window(lens, g, recs, x0, y0, w, h)
    assert win.sum() > 0
    np.testing.assert_allclose(win, full[y0:y0 + h, x0:x0 + w], rtol=1e-12, atol=1e-15)


def test_irs_binning_rules(oracle):
    # irs_functions.cuh:152-184: inclusive border, truncation, clamp, row flip
    h, n = 5.0, 10
    assert oracle.irs_bin(-5.0, -5.0, h, h, n, n) == (n - 1) * n + 0
    assert oracle.irs_bin(5.0, 5.0, h, h, n, n) == 0 * n + (n - 1)
    assert oracle.irs_bin(5.0000001, 0, h, h, n, n) == -1
    assert oracle.irs_bin(0.0, 0.0, h, h, n, n) == (n - 1 - 5) * n + 5
    assert oracle.irs_bin(-4.99, 4.99, h, h, n, n) == 0 * n + 0


def test_irs_counts_rays(oracle, c1_tree):
    T, stars, R = c1_tree
    T.fmm(16)
    d = oracle.derived_params(0.3, 0.3, 0.3, 1.0, "equal", 1.0, 1.0, 1.0, 0.001, 0, 1, 1.37, (0, 0), (10, 10),
                              (50, 50), 9)
    lens = lens_for(oracle, 0.3, 0.3, 0.3, R)
    g = oracle.make_grid(d, (0.0, 0.0), (10.0, 10.0), (50, 50))
    rows = np.arange(0, d.num_ray_threads[1], 5)
    m, got = T.irs_rows(lens, g, rows)
    assert m.sum() == got <= len(rows) * d.num_ray_threads[0]
    tot, mn, sd, got2 = T.irs_rows(lens, g, rows, parities=True)
    np.testing.assert_array_equal(mn + sd, tot)
    np.testing.assert_array_equal(tot, m)
