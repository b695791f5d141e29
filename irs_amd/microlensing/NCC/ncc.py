"""NCC: number of caustic crossings over a source-plane pixel grid.

Drop-in for the NCC class of microlensing/NCC/ncc.py:6-219 over lib_ncc.so:
the same constructor keywords, attributes, ValueError messages, run(), save()
and plots. The caustics are read from <infile_prefix>ccf_caustics.bin, the
file CCF.save() writes.

Every attribute is one entry of lib_ncc's field table; the properties are
generated from that table and the validation rules below, so the class holds
no per-attribute code.
"""

import numpy as np

from . import lib_ncc
from .. import _plotting

# attribute -> (accepts(value), message of the ValueError otherwise)
_RULES = {
    "half_length_y1": (lambda v: v >= 0, "half_length_y1 must be > 0"),
    "half_length_y2": (lambda v: v >= 0, "half_length_y2 must be > 0"),
    "num_pixels_y1": (lambda v: v >= 1, "num_pixels_y1 must be >= 1"),
    "num_pixels_y2": (lambda v: v >= 1, "num_pixels_y2 must be >= 1"),
    "over_sample": (lambda v: v >= 0, "over_sample must be >= 0"),
}


def _attribute(name, kind):
    """Property over get_<name> / set_<name>. Strings travel as UTF-8 bytes,
    write_* flags as ints read back as bools; None leaves the value as is."""
    rule = _RULES.get(name)
    is_flag = name.startswith("write_")

    def fget(self):
        value = getattr(self.lib, "get_" + name)(self.obj)
        if kind == "s":
            return value.decode("utf-8")
        return bool(value) if is_flag else value

    def fset(self, value):
        if value is None:
            return
        if rule is not None and not rule[0](value):
            raise ValueError(rule[1])
        if kind == "s":
            value = value.encode("utf-8")
        elif is_flag:
            value = int(value)
        getattr(self.lib, "set_" + name)(self.obj, value)

    return property(fget, fset)


class NCC(object):
    """Caustic-crossing counts of a square or rectangular source-plane region.

    Keywords (None keeps the library default, as in the reference):
      infile_prefix        where ccf_caustics.bin is read from
      center_y1/2          centre of the map
      half_length_y1/2     half-lengths of the map
      num_pixels_y1/2      pixels along each axis
      over_sample          the map is first counted on 2^over_sample sub-pixels
                           per pixel and axis (4**over_sample times the memory)
      write_maps, write_histograms, outfile_prefix   what save() writes, and where
    """

    def __init__(self, infile_prefix: str = None, center_y1: float = None, center_y2: float = None,
                 half_length_y1: float = None, half_length_y2: float = None, num_pixels_y1: int = None,
                 num_pixels_y2: int = None, over_sample: int = None, write_maps: bool = None,
                 write_histograms: bool = None, outfile_prefix: str = None, verbose: int = 0):
        given = dict(locals())
        self.lib = lib_ncc.lib
        self.obj = self.lib.init()
        self.verbose = verbose
        self.num_caustic_crossings = None
        for name, _ in lib_ncc._FIELDS:
            setattr(self, name, given[name])

    def __del__(self):
        self.close()

    def close(self):
        """Free the library object; the class is unusable afterwards."""
        if getattr(self, "obj", None):
            self.lib.ncc_free(self.obj)
            self.obj = None

    center = property(lambda self: (self.center_y1, self.center_y2))
    half_length = property(lambda self: (self.half_length_y1, self.half_length_y2))
    num_pixels = property(lambda self: (self.num_pixels_y1, self.num_pixels_y2))
    pixel_scales = property(lambda self: tuple(2 * h / n for h, n in zip(self.half_length, self.num_pixels)))
    t_ncc = property(lambda self: self.lib.get_t_ncc(self.obj))
    t_reduce = property(lambda self: self.lib.get_t_reduce(self.obj))

    def run(self):
        """Count the crossings; the int32 map (row 0 at the top) lands in
        num_caustic_crossings."""
        if not self.lib.run(self.obj, self.verbose):
            raise Exception("Error running NCC")
        rows, cols = self.num_pixels_y2, self.num_pixels_y1
        self.num_caustic_crossings = np.array(
            np.ctypeslib.as_array(self.lib.get_num_crossings(self.obj), shape=(rows, cols)))

    def save(self):
        if not self.lib.save(self.obj, self.verbose):
            raise Exception("Error saving NCC")

    def _count_range(self):
        """Half-integer bounds around the smallest and largest count."""
        n = self.num_caustic_crossings
        return float(np.min(n)) - 0.5, float(np.max(n)) + 0.5

    def plot_map(self, fig, ax, **kwargs):
        lo, hi = self._count_range()
        _plotting.image(fig, ax, self.num_caustic_crossings, _plotting.box(self.center, self.half_length),
                        self.half_length[0] / self.half_length[1], "$N_{\\text{microminima}}$",
                        ("$y_1$", "$y_2$"), kwargs, {"vmin": lo, "vmax": hi, "cmap": "viridis"},
                        levels=lambda o: int(o["vmax"] - o["vmin"]))

    def plot_hist(self, fig, ax, bins=None, **kwargs):
        if bins is None:
            lo, hi = self._count_range()
            bins = np.arange(lo, hi + 1, 1)
        _plotting.density(ax, self.num_caustic_crossings, bins, "$N_{\\text{microminima}}$",
                          "p($N_{\\text{microminima}}$)", kwargs)


for _name, _kind in lib_ncc._FIELDS:
    setattr(NCC, _name, _attribute(_name, _kind))
del _name, _kind
