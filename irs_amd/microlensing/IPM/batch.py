"""Many IPM maps in one persistent, graph-captured launch.

IPMBatch gathers IPM objects (float library, mixed precision) that have been
prepared and built, shoots all of them with one HIP graph launch (zeroing,
one persistent kernel over every instance's lattice tiles, histograms) and
returns a 128-bin histogram of log10(magnification) over [-2, 2] per map
(values below 10^-2, zero included, count in the first bin; above 10^2 in
the last). Instance maps are copied to each object's `magnifications` by its
finish(). The first run times candidate tile heights and resident blocks per
CU on a sample of the tiles and keeps the fastest (`tile_h`,
`blocks_per_cu`).

draw_instances() makes the heterogeneous set the batch is benchmarked on:
kappa_tot ~ U(0.1, 0.9), shear ~ U(0, 0.9), smooth fraction ~ U(0, 0.9),
Kroupa stars, square maps with sides drawn from {256, 512, 1024, 2048}, and a
source half-length chosen per instance so that its star field holds a
log-uniform 5e3-1e5 stars. The FMM order is fixed (20 by default): the
automatic order of small fields can fall below the minimum of 3.
"""

import ctypes
import math

import numpy as np

from . import lib_ipm
from .ipm import IPM

HIST_BINS = 128
HIST_RANGE = (-2.0, 2.0)


class IPMBatch:
    def __init__(self, verbose: int = 0):
        self.lib = lib_ipm.lib
        self.obj = self.lib.ipm_batch_create()
        if not self.obj:
            raise Exception("Error creating the IPM batch")
        self.verbose = verbose
        self.instances = []
        self.histograms = None

    def add(self, ipm: IPM) -> int:
        """Add a prepared and built IPM object of the float library in mixed precision."""
        if ipm.is_double:
            raise ValueError("batched maps use the float library (is_double=False)")
        idx = self.lib.ipm_batch_add(self.obj, ipm.obj)
        if idx < 0:
            raise Exception("Error adding an instance to the IPM batch")
        self.instances.append(ipm)
        return idx

    def run(self, copy_maps: bool = True):
        """Shoot every instance; returns the (n, 128) histogram counts."""
        if not self.lib.ipm_batch_shoot(self.obj, self.verbose):
            raise Exception("Error shooting the IPM batch")
        h = np.zeros((len(self.instances), HIST_BINS), dtype=np.uint32)
        if not self.lib.ipm_batch_histograms(self.obj, h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))):
            raise Exception("Error reading the batch histograms")
        self.histograms = h
        if copy_maps:
            for m in self.instances:
                m.finish()
        return h

    @property
    def t_shoot(self):
        """Seconds of the last graph launch (zeroing, shooting, histograms)."""
        return self.lib.get_batch_t_shoot(self.obj)

    @property
    def cells(self):
        return self.lib.get_batch_cells(self.obj)

    @property
    def tile_h(self):
        return self.lib.get_batch_tile_h(self.obj)

    @property
    def blocks_per_cu(self):
        return self.lib.get_batch_blocks_per_cu(self.obj)

    def set_tuning(self, tile_h, blocks_per_cu):
        """Use these tiles instead of autotuning: tile_h lattice rows per tile and
        blocks_per_cu resident blocks of one wave each per CU (12 fill a CU)."""
        if not self.lib.ipm_batch_set_tuning(self.obj, int(tile_h), int(blocks_per_cu)):
            raise ValueError("invalid batch tuning")

    @property
    def t_tune(self):
        return self.lib.get_batch_t_tune(self.obj)

    def close(self):
        if self.obj:
            self.lib.ipm_batch_free(self.obj)
            self.obj = None


def _stars_for(params, half_length):
    m = IPM(**params, half_length_y1=half_length, half_length_y2=half_length)
    m.derive()
    n = m.num_stars
    m.close()
    return n


def draw_instances(n, seed=0, sides=(256, 512, 1024, 2048), star_range=(5.0e3, 1.0e5), num_rays_y=4,
                   kappa_range=(0.1, 0.9), shear_range=(0.0, 0.9), smooth_range=(0.0, 0.9), min_gap=0.05,
                   m_lower=0.08, m_upper=1.5, min_half_length=2.0, expansion_order=20):
    """Parameter dictionaries for n IPM instances (see the module docstring).
    Draws with |1 - kappa - shear| or |1 - kappa + shear| below min_gap are
    rejected (the lattice grows as 1/|1 - kappa -+ shear|), and so are draws
    whose star target is reached below min_half_length Einstein radii (there
    the light-loss margin, not the map, sizes the field and the lattice)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        kappa = float(rng.uniform(*kappa_range))
        shear = float(rng.uniform(*shear_range))
        smooth = float(rng.uniform(*smooth_range))
        if abs(1 - kappa - shear) < min_gap or abs(1 - kappa + shear) < min_gap:
            continue
        side = int(rng.choice(sides))
        target = math.exp(rng.uniform(math.log(star_range[0]), math.log(star_range[1])))
        base = dict(kappa_tot=kappa, shear=shear, kappa_star=(1 - smooth) * kappa, mass_function="kroupa",
                    m_lower=m_lower, m_upper=m_upper, num_pixels_y1=side, num_pixels_y2=side, num_rays_y=num_rays_y,
                    random_seed=int(rng.integers(1, 2 ** 31 - 1)), write_stars=False, write_maps=False,
                    write_histograms=False, precision=1, expansion_order=expansion_order)
        lo, hi = min_half_length, 2.0 * min_half_length
        if _stars_for(base, lo) > target:
            continue
        while _stars_for(base, hi) < target:
            hi *= 2
            if hi > 1e4:
                break
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if _stars_for(base, mid) < target:
                lo = mid
            else:
                hi = mid
        base.update(half_length_y1=hi, half_length_y2=hi)
        out.append(base)
    return out
