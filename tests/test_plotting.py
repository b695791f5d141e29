"""The plot methods of the IPM, CCF and NCC classes on synthetic results (no
device: the arrays are set by hand), checking what each figure shows against
the behaviour of microlensing/IPM/ipm.py:442-473, microlensing/CCF/ccf.py:344-361
and microlensing/NCC/ncc.py:187-219."""

import numpy as np
import pytest

mpl = pytest.importorskip("matplotlib")
mpl.use("Agg")
import matplotlib.pyplot as plt  # noqa: E402


def test_ipm_plots():
    from irs_amd.microlensing.IPM.ipm import IPM
    m = IPM(kappa_tot=0.3, shear=0.3, kappa_star=0.3, theta_star=2.0, half_length_y1=4.0, half_length_y2=2.0,
            center_y1=1.0, num_pixels_y1=20, num_pixels_y2=10, is_double=True)
    rng = np.random.default_rng(0)
    m.magnifications = rng.uniform(0.5, 20.0, size=(10, 20))
    fig, ax = plt.subplots()
    m.plot_map(fig, ax)
    img = ax.get_images()[0]
    dm = -2.5 * np.log10(m.magnifications / abs(m.mu_ave))
    np.testing.assert_allclose(img.get_array(), dm)
    assert img.get_extent() == [(1 - 4) / 2, (1 + 4) / 2, -1.0, 1.0]
    assert img.get_clim() == (np.percentile(dm, 2.5), np.percentile(dm, 97.5))
    assert img.get_cmap().name == "viridis_r"
    assert ax.get_aspect() == 2.0
    fig2, ax2 = plt.subplots()
    m.plot_map(fig2, ax2, vmin=-1.0, cmap="magma")
    assert ax2.get_images()[0].get_clim()[0] == -1.0 and ax2.get_images()[0].get_cmap().name == "magma"
    fig3, ax3 = plt.subplots()
    m.plot_hist(fig3, ax3)
    heights = np.array([p.get_height() for p in ax3.patches])
    widths = np.array([p.get_width() for p in ax3.patches])
    assert np.allclose(widths, 0.01) and abs((heights * widths).sum() - 1) < 1e-9
    assert ax3.xaxis_inverted()
    plt.close("all")
    m.close()


def test_ccf_plots():
    from irs_amd.microlensing.CCF.ccf import CCF
    c = CCF(kappa_tot=0.3, shear=0.1, kappa_star=0.3)
    t = np.linspace(0, 2 * np.pi, 9)
    ring = np.stack([np.cos(t), np.sin(t)], axis=-1)
    c.critical_curves = np.stack([ring, 2 * ring])
    c.caustics = np.stack([ring + 1, ring - 1, 3 * ring])
    fig, ax = plt.subplots()
    c.plot_critical_curves(fig, ax, color="k")
    assert len(ax.lines) == 2
    np.testing.assert_allclose(ax.lines[1].get_xydata(), 2 * ring)
    assert "x_1" in ax.get_xlabel()
    c.plot_caustics(fig, ax)
    assert len(ax.lines) == 5 and "y_2" in ax.get_ylabel()
    plt.close("all")
    c.close()


def test_ncc_plots():
    from irs_amd.microlensing.NCC.ncc import NCC
    n = NCC(center_y1=1.0, center_y2=-1.0, half_length_y1=3.0, half_length_y2=1.5, num_pixels_y1=6,
            num_pixels_y2=3)
    n.num_caustic_crossings = np.array([[0, 2, 2, 4, 4, 6], [0, 0, 2, 2, 4, 4], [0, 0, 0, 2, 2, 2]], dtype=np.int32)
    fig, ax = plt.subplots()
    n.plot_map(fig, ax)
    img = ax.get_images()[0]
    assert img.get_clim() == (-0.5, 6.5) and img.get_cmap().N == 7
    assert img.get_extent() == [-2.0, 4.0, -2.5, 0.5] and ax.get_aspect() == 2.0
    fig2, ax2 = plt.subplots()
    n.plot_hist(fig2, ax2)
    edges = [p.get_x() for p in ax2.patches]
    heights = [p.get_height() for p in ax2.patches]
    # one unit-wide bin per count, centred on it
    assert np.allclose(edges, np.arange(-0.5, 6.5, 1.0))
    assert np.allclose(heights, [6 / 18, 0, 7 / 18, 0, 4 / 18, 0, 1 / 18])
    plt.close("all")
    n.close()
