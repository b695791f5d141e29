"""Readers and writers for the files the map makers produce.

Same functions, arguments and byte layouts as the reference's
microlensing/Util/util.py:5-125:

* parameter files: whitespace-separated "key value" lines
  (ipm.cuh:1141-1206 writes them);
* star files: int32 num_stars, int32 rectangular, dtype corner[2],
  dtype theta_star, then num_stars x (x1, x2, mass) in dtype
  (star.cuh:498-534);
* 2D arrays (maps): int32 nrows, int32 ncols, then nrows x ncols values
  (array_functions.cuh:280-316);
* histograms: "value count" text lines (array_functions.cuh:318-349).
"""

import numpy as np

_DTYPES = {np.dtype(np.float32): np.float32, np.dtype(np.float64): np.float64}


def _parse_value(text):
    try:
        v = float(text)
    except ValueError:
        return text
    return int(v) if v == int(v) else v


def read_params(fname: str):
    """Dictionary of the "key value" lines of a parameter file. Numbers
    become int when integral, else float; anything else stays a string."""
    params = {}
    with open(fname) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            key, val = parts
            params[key] = _parse_value(val)
    return params


def read_stars(fname: str, dtype=np.float32):
    """(stars[n, 3], rectangular, corner[2], theta_star) from a star file."""
    dtype = np.dtype(dtype)
    with open(fname, "rb") as fh:
        head = np.fromfile(fh, dtype=np.int32, count=2)
        nstars, rectangular = int(head[0]), bool(head[1])
        corner = np.fromfile(fh, dtype=dtype, count=2)
        theta_star = np.fromfile(fh, dtype=dtype, count=1)[0]
        stars = np.fromfile(fh, dtype=dtype).reshape(nstars, 3)
    return stars, rectangular, corner, theta_star


def write_stars(fname: str, nstars: int, rectangular: bool, corner, theta_star: float, stars, dtype=np.float32):
    """Write a star file; corner is (x1, x2) of a rectangular field or
    (radius, 0) of a circular one."""
    if not fname.endswith(".bin"):
        raise ValueError("fname must be a .bin file")
    if np.dtype(dtype) not in _DTYPES:
        raise ValueError("dtype must be np.float32 or np.float64")
    dtype = np.dtype(dtype)
    stars = np.asarray(stars, dtype=dtype).reshape(-1)
    with open(fname, "wb") as fh:
        np.array([nstars, 1 if rectangular else 0], dtype=np.int32).tofile(fh)
        np.asarray(corner, dtype=dtype)[:2].tofile(fh)
        np.array([theta_star], dtype=dtype).tofile(fh)
        stars.tofile(fh)


def read_array(fname: str, dtype, is_complex: bool = False):
    """A 2D array file as an (nrows, ncols) array, or (nrows, ncols, 2) for
    complex values."""
    with open(fname, "rb") as fh:
        nrows, ncols = (int(v) for v in np.fromfile(fh, dtype=np.int32, count=2))
        dat = np.fromfile(fh, dtype=dtype)
    return dat.reshape(nrows, ncols, 2) if is_complex else dat.reshape(nrows, ncols)


def write_array(fname: str, arr, dtype=None):
    """Write a 2D array file (the layout read_array reads)."""
    arr = np.asarray(arr if dtype is None else np.asarray(arr, dtype=dtype))
    if arr.ndim != 2:
        raise ValueError("arr must be two-dimensional")
    with open(fname, "wb") as fh:
        np.array(arr.shape, dtype=np.int32).tofile(fh)
        np.ascontiguousarray(arr).tofile(fh)


def read_hist(fname: str):
    """Histogram file as an int32 array of (value, num_pixels) rows (as
    np.loadtxt returns it: one row comes back one-dimensional)."""
    return np.loadtxt(fname, dtype=np.int32)
