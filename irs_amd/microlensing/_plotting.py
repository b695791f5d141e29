"""Matplotlib helpers shared by the map makers' plot_* methods.

The plot methods of the reference (microlensing/IPM/ipm.py:442-473,
microlensing/CCF/ccf.py:344-361, microlensing/NCC/ncc.py:187-219) draw three
kinds of figure: a map as an image over its source-plane extent, the density
histogram of a map's values, and a family of curves. Each kind is one function
here; the classes only say what to draw. Defaults a caller passes in kwargs win
over the ones given here. matplotlib is imported on first use, so the package
imports without it.
"""

import numpy as np

Y_LABELS = ("$y_1 / \\theta_★$", "$y_2 / \\theta_★$")
X_LABELS = ("$x_1 / \\theta_★$", "$x_2 / \\theta_★$")


def box(center, half_length, unit=1.0):
    """imshow extent [left, right, bottom, top] of a centred box, in `unit`."""
    (c1, c2), (h1, h2) = center, half_length
    return [(c1 - h1) / unit, (c1 + h1) / unit, (c2 - h2) / unit, (c2 + h2) / unit]


def image(fig, ax, data, extent, aspect, colorbar_label, axis_labels, kwargs, defaults, levels=None,
          flip_colorbar=False):
    """Draw `data` (row 0 at the top) over `extent`.

    `defaults` fills the imshow keywords the caller left out. With `levels`,
    the colour map is cut into that many discrete colours (integer maps)."""
    import matplotlib.pyplot as plt
    opts = dict(defaults)
    opts.update(kwargs)
    if levels is not None:
        opts["cmap"] = plt.get_cmap(opts["cmap"], levels(opts))
    img = ax.imshow(data, extent=extent, **opts)
    bar = fig.colorbar(img, label=colorbar_label)
    if flip_colorbar:
        bar.ax.invert_yaxis()
    ax.set_xlabel(axis_labels[0])
    ax.set_ylabel(axis_labels[1])
    ax.set_aspect(aspect)
    return img


def density(ax, values, bins, xlabel, ylabel, kwargs, flip_x=False):
    """Normalised histogram of every value of `values`."""
    out = ax.hist(np.ravel(values), bins=bins, density=True, **kwargs)
    ax.set_xlabel(xlabel)
    ax.set_ylabel(ylabel)
    if flip_x:
        ax.invert_xaxis()
    return out


def curves(ax, points, axis_labels, kwargs):
    """One line per row of `points` (rows x samples x 2 of (re, im))."""
    pts = np.asarray(points)
    lines = ax.plot(pts[..., 0].T, pts[..., 1].T, **kwargs)
    ax.set_xlabel(axis_labels[0])
    ax.set_ylabel(axis_labels[1])
    return lines
