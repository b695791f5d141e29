"""Critical curves and caustics (ccf_functions.cuh, ccf.cuh).

CPU: the oracle (oracle/ccf.py) is pinned by the single-star closed form of
the roots and caustics, by finite differences of the smooth-matter deflection
(its first and second derivatives), and by a finite-difference construction of
the magnification length scale from the lens equation alone. lib_ccf.so
validates its parameters before touching the GPU.

GPU: the library's roots, caustics and length scales match the closed form and
the oracle (direct sums; the library's tree/FMM carries an alpha_error of
1e-7 Einstein radii), and every root of the rectangular and exact
smooth-matter variants solves the critical-curve equation.
"""

import ctypes
import os

import numpy as np
import pytest

from oracle import ccf as OC

TOL_ROOT = 1e-5   # Einstein radii, library (tree/FMM) vs oracle (direct sums)


def single_star_roots(z0, m, kappa, kappastar, gamma, theta, phi):
    """z0 +- sqrt(theta^2 m / ((1 - kappa + kappa_star) e^{-i phi} - gamma)) (approximate circular sheet)."""
    D = 1 - kappa + kappastar
    r = np.sqrt(theta ** 2 * m / (D * np.exp(-1j * phi) - gamma))
    return np.array([z0 + r, z0 - r])


def match_sets(a, b):
    """Largest distance from each point of a to the nearest of b, and back."""
    d = np.abs(a[:, None] - b[None, :])
    return max(d.min(1).max(), d.min(0).max())


def test_single_star_roots_and_caustics_closed_form(oracle):
    z0, m = 0.3 - 0.2j, 1.0
    kappa, kappastar, gamma, theta = 0.3, 0.3, 0.4, 1.0
    prob = OC.Problem([[z0.real, z0.imag, m]], kappa, gamma, theta, kappastar, (2.0, 0.0))
    # two branches: at the one-branch centre phase pi the roots lie on the line through the star
    # perpendicular to the starting offsets (+-1), which the iteration cannot leave
    roots, phis, err = OC.solve(prob, 8, 2)
    assert roots.shape == (2, 10) and err < 1e-9
    for col, phi in enumerate(phis):
        assert match_sets(roots[:, col], single_star_roots(z0, m, kappa, kappastar, gamma, theta, phi)) < 1e-10
    z = roots.ravel()
    w = (1 - kappa + kappastar) * z + gamma * np.conj(z) - theta ** 2 * m / np.conj(z - z0)
    assert np.abs(prob.source(z) - w).max() < 1e-10


@pytest.mark.parametrize("rectangular,approx", [(1, 1), (1, 0), (0, 0)])
def test_smooth_derivatives_match_finite_differences(oracle, rectangular, approx):
    kappa, gamma, kappastar, corner, ts = 0.4, 0.2, 0.3, (3.0, 2.0), 21
    lens = oracle.make_lens(kappa, gamma, 1.0, kappastar, corner, rectangular, approx, ts)
    empty = np.zeros((0, 3))

    def alpha(z):
        z = np.asarray(z, dtype=complex)
        return (1 - kappa) * z + gamma * np.conj(z) - oracle.deflect_direct(empty, lens, z)

    z = np.array([0.3 + 0.2j, -1.1 + 0.7j, 0.9 - 1.3j])
    if not rectangular:
        # outside the circle of radius |corner|: inside it the reference keeps the outside form of
        # d alpha_smooth / dzbar (alpha_smooth.cuh:228-231), which is not the derivative there
        z = np.array([5 + 1j, -4 + 3j, 2 - 6j])
    h = 1e-5
    ax = (alpha(z + h) - alpha(z - h)) / (2 * h)
    ay = (alpha(z + 1j * h) - alpha(z - 1j * h)) / (2 * h)
    ddz, d1, d2 = OC.smooth_terms(z, kappastar, rectangular, approx, ts, corner)
    assert np.allclose((ax - 1j * ay) / 2, ddz, rtol=0, atol=1e-7)
    assert np.allclose((ax + 1j * ay) / 2, d1, rtol=0, atol=1e-7)
    _, d1p, _ = OC.smooth_terms(z + h, kappastar, rectangular, approx, ts, corner)
    _, d1m, _ = OC.smooth_terms(z - h, kappastar, rectangular, approx, ts, corner)
    _, d1pi, _ = OC.smooth_terms(z + 1j * h, kappastar, rectangular, approx, ts, corner)
    _, d1mi, _ = OC.smooth_terms(z - 1j * h, kappastar, rectangular, approx, ts, corner)
    d2_fd = ((d1p - d1m) + 1j * (d1pi - d1mi)) / (4 * h)
    assert np.allclose(d2_fd, d2, rtol=0, atol=1e-6 * max(1, np.abs(d2).max()))


def test_mu_length_scale_from_the_lens_equation(oracle):
    """1 / (2 |grad det J| |dw along the unit tangent|) from finite differences of w alone."""
    stars = np.array([[0.0, 0.0, 1.0], [1.3, 0.4, 0.6], [-0.8, 1.1, 1.4]])
    prob = OC.Problem(stars, 0.3, 0.25, 1.0, 0.3, (3.0, 0.0))
    roots, _, _ = OC.solve(prob, 8, 2)
    z = roots[:, ::3].ravel()

    def jac_det(p, e=1e-6):
        wx = (prob.source(p + e) - prob.source(p - e)) / (2 * e)
        wy = (prob.source(p + 1j * e) - prob.source(p - 1j * e)) / (2 * e)
        return wx.real * wy.imag - wx.imag * wy.real

    e = 1e-4
    grad = ((jac_det(z + e) - jac_det(z - e)) + 1j * (jac_det(z + 1j * e) - jac_det(z - 1j * e))) / (2 * e)
    u = 1j * grad / np.abs(grad)
    dw = (prob.source(z + e * u) - prob.source(z - e * u)) / (2 * e)
    d0 = 1 / (2 * np.abs(grad) * np.abs(dw))
    assert np.allclose(prob.mu_length_scale(z), d0, rtol=1e-4)


def test_oracle_roots_are_critical(oracle):
    rng = np.random.default_rng(3)
    n = 12
    r = 3.0 * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    stars = np.column_stack([r * np.cos(a), r * np.sin(a), np.ones(n)])
    prob = OC.Problem(stars, 0.45, 0.2, 1.0, 0.4, (3.0, 0.0))
    roots, phis, err = OC.solve(prob, 16, 2)
    assert roots.shape == (2 * n, 18) and err < 1e-8
    assert np.abs(prob.det_jacobian(roots)).max() < 1e-7
    for col in range(roots.shape[1]):
        d = np.abs(roots[:, col, None] - roots[None, :, col])
        np.fill_diagonal(d, np.inf)
        assert d.min() > 1e-6


def test_library_validates_before_the_device(capfd):
    from irs_amd.microlensing.CCF import lib_ccf
    lib = lib_ccf.lib
    for setter, val, msg in (("set_num_phi", 7, "num_phi must be an even integer > 0"),
                             ("set_num_branches", 3, "num_phi must be a multiple of 2 * num_branches"),
                             ("set_safety_scale", 1.0, "safety_scale must be >= 1.1"),
                             ("set_num_stars", 0, "num_stars must be an integer > 0"),
                             ("set_mass_function", b"lognormal", "mass_function must be equal")):
        obj = lib.init()
        getattr(lib, setter)(obj, val)
        assert not lib.run(obj, 0)
        assert msg in capfd.readouterr().err
        lib.ccf_free(obj)


def test_library_defaults_match_reference():
    """ccf.cuh:42-62."""
    from irs_amd.microlensing.CCF import lib_ccf
    lib = lib_ccf.lib
    obj = lib.init()
    assert (lib.get_kappa_tot(obj), lib.get_shear(obj), lib.get_kappa_star(obj), lib.get_theta_star(obj)) == (0.3, 0.3, 0.27, 1.0)
    assert lib.get_mass_function(obj) == b"equal" and lib.get_num_stars(obj) == 137
    assert (lib.get_num_phi(obj), lib.get_num_branches(obj), lib.get_safety_scale(obj)) == (100, 1, 1.37)
    assert (lib.get_write_stars(obj), lib.get_write_critical_curves(obj), lib.get_write_caustics(obj),
            lib.get_write_mu_length_scales(obj)) == (1, 1, 1, 0)
    lib.ccf_free(obj)


def test_python_mirror_value_errors():
    from irs_amd.microlensing.CCF.ccf import CCF
    with pytest.raises(ValueError, match="even"):
        CCF(kappa_tot=0.3, shear=0.2, kappa_star=0.3, num_phi=21)
    with pytest.raises(ValueError, match="multiple of 2"):
        CCF(kappa_tot=0.3, shear=0.2, kappa_star=0.3, num_phi=20, num_branches=3)
    with pytest.raises(ValueError, match="kappa_star must be <= kappa_tot"):
        CCF(kappa_tot=0.3, shear=0.2, kappa_star=0.4)
    with pytest.raises(ValueError, match="safety_scale"):
        CCF(kappa_tot=0.3, shear=0.2, kappa_star=0.3, safety_scale=1.0)


def run_ccf(**kw):
    from irs_amd.microlensing.CCF.ccf import CCF
    c = CCF(write_length_scales=True, **kw)
    c.run()
    return c


@pytest.mark.gpu
def test_gpu_single_star_closed_form():
    kappa, kappastar, gamma = 0.3, 0.25, 0.4
    c = run_ccf(kappa_tot=kappa, shear=gamma, kappa_star=kappastar, num_stars=1, num_phi=12, num_branches=2,
                random_seed=5, write_stars=False)
    s = c.stars[0]
    z0 = s[0] + 1j * s[1]
    cc = c.critical_curves[..., 0] + 1j * c.critical_curves[..., 1]
    assert cc.shape == (4, 7)
    for b in range(2):
        for k in range(7):
            phi = np.pi / 2 + np.pi * b + (k - 3) * 2 * np.pi / 12
            assert match_sets(cc[b::2, k], single_star_roots(z0, s[2], kappa, kappastar, gamma, 1.0, phi)) < 1e-9
    w = c.caustics[..., 0] + 1j * c.caustics[..., 1]
    expect = (1 - kappa + kappastar) * cc + gamma * np.conj(cc) - s[2] / np.conj(cc - z0)
    assert np.abs(w - expect).max() < 1e-9
    assert c.max_error < 1e-9


@pytest.mark.gpu
def test_gpu_matches_oracle(oracle):
    kappa, kappastar, gamma, n, seed = 0.45, 0.4, 0.3, 64, 7
    c = run_ccf(kappa_tot=kappa, shear=gamma, kappa_star=kappastar, num_stars=n, num_phi=12, num_branches=1,
                random_seed=seed, write_stars=False)
    # the same Philox field in generation order (equal masses: no corner rescale)
    q = np.array([1 / abs(1 - kappa + gamma), 1 / abs(1 - kappa - gamma)])
    corner = q / np.hypot(*q) * np.sqrt(n / kappastar)
    stars = oracle.generate_stars(n, 0, tuple(corner), "equal", 1.0, 1.0, 1.0, seed)
    assert np.allclose(np.sort(stars[:, 0]), np.sort(c.stars[:, 0]), atol=1e-12)
    prob = OC.Problem(stars, kappa, gamma, 1.0, kappastar, c.corner)
    roots, phis, err = OC.solve(prob, 12, 1)
    cc = c.critical_curves[..., 0] + 1j * c.critical_curves[..., 1]
    assert cc.shape == roots.shape == (2 * n, 13)
    for col in range(13):
        assert match_sets(cc[:, col], roots[:, col]) < TOL_ROOT
    # every library root solves the equation by direct sums, and its caustic point is the lens equation there
    err_gpu = prob.inv_mu_error(cc, phis[None, :])
    assert err_gpu.max() < 1e-6
    w = c.caustics[..., 0] + 1j * c.caustics[..., 1]
    assert np.abs(w - prob.source(cc)).max() < 1e-6
    assert np.allclose(c.mu_length_scales, prob.mu_length_scale(cc), rtol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("rectangular,approx", [(1, 1), (0, 0), (1, 0)])
def test_gpu_smooth_variants_solve_the_equation(oracle, rectangular, approx):
    kappa, kappastar, gamma = 0.4, 0.3, 0.25
    c = run_ccf(kappa_tot=kappa, shear=gamma, kappa_star=kappastar, num_stars=24, num_phi=8, num_branches=2,
                rectangular=rectangular, approx=approx, random_seed=11, write_stars=False)
    stars = c.stars
    lib = c.lib
    # taylor_smooth is not exported; the oracle recomputes it as ccf.cuh:410-431 does
    corner = np.array(c.corner)
    ts = 1
    if rectangular:
        cabs, ph = np.hypot(*corner), np.arctan2(corner[1], corner[0])
        s = 1.37
        while kappastar / np.pi * 4 / (ts + 1) * cabs * (s + 1) / (s - 1) * (1 / s) ** (ts + 1) > 1e-7 and ts <= 101:
            ts += 2
        while (np.fmod(ph * (ts - 1), np.pi) < 0.1 * np.pi or np.fmod(ph * (ts - 1), np.pi) > 0.9 * np.pi) and ts <= 101:
            ts += 2
    nroots = OC.num_roots(24, rectangular, approx, ts)
    assert c.num_roots == nroots
    prob = OC.Problem(stars, kappa, gamma, 1.0, kappastar, tuple(corner), rectangular, approx, ts)
    cc = c.critical_curves[..., 0] + 1j * c.critical_curves[..., 1]
    assert cc.shape == (nroots * 2, 5)
    # branch b of root r is row 2 r + b; its phases pi/2 + pi b + (k - 2) pi/4
    phis = np.array([[np.pi / 2 + np.pi * b + (k - 2) * np.pi / 4 for k in range(5)] for b in range(2)])
    err = prob.inv_mu_error(cc, np.tile(phis, (nroots, 1)))
    assert np.nanmax(err) < 1e-6
    assert c.max_error < 1e-6
    del lib


@pytest.mark.gpu
def test_gpu_saved_files_read_back(tmp_path):
    from irs_amd.microlensing.Util import util
    prefix = str(tmp_path) + "/"
    c = run_ccf(kappa_tot=0.4, shear=0.3, kappa_star=0.35, num_stars=10, num_phi=10, num_branches=1, random_seed=2,
                write_stars=True, write_critical_curves=True, write_caustics=True, outfile_prefix=prefix)
    c.save()
    info = open(prefix + "ccf_parameter_info.txt").read()
    for key in ("kappa_tot", "num_roots", "num_phi", "max_error_1/mu", "t_ccs", "rad"):
        assert f"\n{key} " in "\n" + info
    for name, arr in (("ccf_ccs.bin", c.critical_curves), ("ccf_caustics.bin", c.caustics)):
        raw = np.fromfile(prefix + name, dtype=np.int32, count=2)
        assert tuple(raw) == arr.shape[:2]
        vals = np.fromfile(prefix + name, dtype=np.float64, offset=8).reshape(arr.shape)
        assert np.array_equal(vals, arr)
    vals = np.fromfile(prefix + "ccf_mu_length_scales.bin", dtype=np.float64, offset=8)
    assert np.array_equal(vals, c.mu_length_scales.ravel())
    stars, rect, corner, theta = util.read_stars(prefix + "ccf_stars.bin", np.float64)
    assert np.array_equal(stars, c.stars)
