"""FP64 restatement of the number of caustic crossings -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of irs_amd/csrc/ncc.hip. The
product never imports it.

It follows ncc_functions.cuh and ncc.cuh of weisluke/irs as literally as
Python allows, with plain IEEE double operations:

* segment selection and clipping: point_in_region, corrected_point and the
  three line tests of find_num_caustic_crossings_kernel
  (ncc_functions.cuh:43-155, 196-236);
* point_to_pixel (array_functions.cuh:70-75);
* for every pixel-centre row a segment crosses, -1 (upwards) or +1
  (downwards) added to EVERY pixel left of the crossing, as the reference's
  per-pixel loop does (here one array slice per crossing,
  ncc_functions.cuh:238-301);
* the over-sampled grid reduced by repeated 2 x 2 maxima and the reference's
  column and row shifts (ncc_functions.cuh:345-411, ncc.cuh:287-316).

The line intersection is evaluated either as the direct product
(y - y1) dx / dy (the default, what the library does) or, with
logspace=True, exactly as the reference writes it:
x1 + sgn(y - y1) sgn(dy) sgn(dx) exp(log|y - y1| - log|dy| + log|dx|).

Pinned by tests/test_ncc.py: for closed clockwise curves, every count equals
the winding number of the curves about the (sub)pixel centre, computed
independently by summing the angles the curve turns through; and the maps of
the two intersection forms are identical on the test curves.
"""

import math

import numpy as np

INT_MAX = 2 ** 31 - 1
INT_MIN = -2 ** 31


def sgn(v):
    """ncc_functions.cuh:24-29."""
    if v < 0:
        return -1.0
    if v > 0:
        return 1.0
    return 0.0


def to_int(v):
    """static_cast<int> as the GPU converts: truncation, NaN -> 0, saturation."""
    if v != v:
        return 0
    if v >= 2147483647.0:
        return INT_MAX
    if v <= -2147483648.0:
        return INT_MIN
    return int(v)


def _log_abs(v):
    return math.log(abs(v)) if v != 0 else -math.inf


def _exp(v):
    if v != v:
        return math.nan
    try:
        return math.exp(v)
    except OverflowError:
        return math.inf


def _mul(a, b):
    """a * b with IEEE semantics (0 * inf = NaN), which Python's float ops already follow."""
    return a * b


def line_x_at(y, p1, p2, logspace=False):
    """get_x_intersection (ncc_functions.cuh:58-84)."""
    dx = p2[0] - p1[0]
    if dx == 0:
        return p1[0]
    dy = p2[1] - p1[1]
    if logspace:
        log_t = _log_abs(y - p1[1]) - _log_abs(dy)
        return p1[0] + _mul(sgn(y - p1[1]) * sgn(dy) * sgn(dx), _exp(log_t + _log_abs(dx)))
    if dy == 0:
        return math.nan
    return p1[0] + (y - p1[1]) * dx / dy


def line_y_at(x, p1, p2, logspace=False):
    """get_y_intersection (ncc_functions.cuh:85-110)."""
    dy = p2[1] - p1[1]
    if dy == 0:
        return p1[1]
    dx = p2[0] - p1[0]
    if logspace:
        log_t = _log_abs(x - p1[0]) - _log_abs(dx)
        return p1[1] + _mul(sgn(x - p1[0]) * sgn(dx) * sgn(dy), _exp(log_t + _log_abs(dy)))
    if dx == 0:
        return math.nan
    return p1[1] + (x - p1[0]) * dy / dx


def in_region(p, hl):
    """point_in_region (ncc_functions.cuh:43-50): x > -hl1 and |y| < hl2."""
    return p[0] > -hl[0] and abs(p[1]) < hl[1]


def corrected(p1, p2, hl, npix, logspace=False):
    """corrected_point (ncc_functions.cuh:124-155)."""
    x, y = p1
    if x <= -hl[0]:
        x = -hl[0] + 0.01 * 2 * hl[0] / float(npix[0])
        y = line_y_at(x, p1, p2, logspace)
    if abs(y) >= hl[1]:
        y = sgn(y) * (hl[1] - 0.01 * 2 * hl[1] / float(npix[1]))
        x = line_x_at(y, p1, p2, logspace)
    return (x, y)


def to_pixel(p, hl, npix):
    """point_to_pixel (array_functions.cuh:70-75)."""
    return ((p[0] + hl[0]) * float(npix[0]) / (2 * hl[0]), (p[1] + hl[1]) * float(npix[1]) / (2 * hl[1]))


def segment_crossings(p0, p1, hl, npix, logspace=False):
    """(row, last column, +-1) of every row the segment crosses, on a grid of
    npix = (columns, rows), row 0 at the top; a last column < 0 adds nothing
    (find_num_caustic_crossings_kernel, ncc_functions.cuh:196-301)."""
    if not (in_region(p0, hl) or in_region(p1, hl)):
        if not (line_x_at(hl[1], p0, p1, logspace) >= -hl[0] or line_x_at(-hl[1], p0, p1, logspace) >= -hl[0]
                or abs(line_y_at(-hl[0], p0, p1, logspace)) <= hl[1]):
            return []
    a = corrected(p0, p1, hl, npix, logspace)
    b = corrected(p1, a, hl, npix, logspace)
    a = to_pixel(a, hl, npix)
    b = to_pixel(b, hl, npix)
    out = []
    if a[1] < b[1]:
        for k in range(to_int(a[1] + 0.5), to_int(b[1] + 0.5)):
            x1 = line_x_at(k + 0.5, a, b, logspace)
            col = min(to_int(x1 + 0.5) - 1, npix[0] - 1)
            out.append((npix[1] - 1 - k, col, -1))
    elif a[1] > b[1]:
        for k in range(to_int(a[1] + 0.5), to_int(b[1] + 0.5), -1):
            x1 = line_x_at(k - 0.5, a, b, logspace)
            col = min(to_int(x1 + 0.5) - 1, npix[0] - 1)
            out.append((npix[1] - 1 - (k - 1), col, 1))
    return out


def fine_counts(caustics, center, half_length, npix, logspace=False, stats=None):
    """Counts on the over-sampled grid npix = (columns, rows), as the
    reference's pixel-by-pixel atomics leave them. caustics: complex array
    (rows, points) or real (rows, points, 2). stats, if given, receives the
    number of crossings and of the per-pixel additions the reference makes."""
    c = np.asarray(caustics)
    if not np.iscomplexobj(c):
        c = c[..., 0] + 1j * c[..., 1]
    num = np.zeros((npix[1], npix[0]), dtype=np.int64)
    hl = (float(half_length[0]), float(half_length[1]))
    for row in c:
        pts = [(float(z.real) - center[0], float(z.imag) - center[1]) for z in row]
        for j in range(len(pts) - 1):
            for r, col, d in segment_crossings(pts[j], pts[j + 1], hl, npix, logspace):
                # the reference's loop over l = 0..col with its flat-index bounds check:
                # a row inside the grid takes every l, a row outside none
                if col >= 0 and 0 <= r < npix[1]:
                    num[r, :col + 1] += d
                    if stats is not None:
                        stats["crossings"] = stats.get("crossings", 0) + 1
                        stats["pixel_updates"] = stats.get("pixel_updates", 0) + col + 1
    return num


def reduce_like_reference(num, num_pixels, over_sample):
    """ncc.cuh:287-316 with ncc_functions.cuh:345-411: over_sample rounds of a
    2 x 2 maximum written to the block's first entry, then the columns and rows
    shifted into place, on the flat array."""
    flat = num.ravel().copy()
    n1, n2 = num_pixels[0] << over_sample, num_pixels[1] << over_sample
    for _ in range(over_sample):
        n1 >>= 1
        n2 >>= 1
        wide = flat[:4 * n1 * n2].reshape(2 * n2, 2 * n1)
        # reduce_pix_array_kernel: the 2 x 2 maximum into the block's first entry
        wide[0::2, 0::2] = np.maximum(np.maximum(wide[0::2, 0::2], wide[0::2, 1::2]),
                                      np.maximum(wide[1::2, 0::2], wide[1::2, 1::2]))
        # shift_pix_column_kernel for columns 1..n1-1 in turn: no column is read after
        # it has been written, so the launches act as one gather
        wide[0::2, 1:n1] = wide[0::2, 2:2 * n1:2].copy()
        # shift_pix_row_kernel for rows 1..n2-1 in turn, likewise a gather
        rows = np.arange(1, n2)
        src = (2 * rows * 2 * n1)[:, None] + np.arange(n1)[None, :]
        dst = (rows * n1)[:, None] + np.arange(n1)[None, :]
        flat[dst] = flat[src].copy()
    return flat[: num_pixels[0] * num_pixels[1]].reshape(num_pixels[1], num_pixels[0])


def ncc_map(caustics, center=(0.0, 0.0), half_length=(5.0, 5.0), num_pixels=(1000, 1000), over_sample=2,
            logspace=False):
    """The number of caustic crossings map (num_pixels[1] rows of
    num_pixels[0], row 0 at the top), as NCC.run() computes it."""
    npix = (num_pixels[0] << over_sample, num_pixels[1] << over_sample)
    fine = fine_counts(caustics, center, half_length, npix, logspace)
    return reduce_like_reference(fine, num_pixels, over_sample)


def histogram(counts):
    """ncc.cuh:332-370: (minimum, counts of each value from the minimum to the maximum)."""
    lo, hi = int(counts.min()), int(counts.max())
    return lo, np.bincount((counts - lo).ravel(), minlength=hi - lo + 1)


def winding_numbers(curves, xs, ys):
    """Clockwise winding number of closed polygonal curves about the points
    (xs, ys), by the sum of the angles each edge subtends (independent of the
    crossing rule above). curves: list of complex arrays of vertices, first
    vertex not repeated."""
    P = xs + 1j * ys
    total = np.zeros(P.shape)
    for c in curves:
        c = np.asarray(c, dtype=complex)
        for k in range(len(c)):
            a, b = c[k] - P, c[(k + 1) % len(c)] - P
            total += np.angle(b / a)
    return np.rint(-total / (2 * np.pi)).astype(np.int64)
