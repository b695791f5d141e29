"""HistogramSurrogate: lens parameters -> 128-bin log10 magnification histogram.

The network (6 -> 256 -> 256 -> 256 -> 128, SiLU, softmax) is trained and run
on the GPU by lib_surrogate.so in BF16 with FP32 accumulation. Inputs are the
lens parameters of draw_instances / ipm_batch: kappa_tot, shear, smooth
fraction, the high-mass slope of the mass function, m_lower and m_upper,
standardised with the training set's mean and spread. Targets are the
normalised histograms of ipm_batch_histograms (128 bins of log10 magnification
over [-2, 2]).

Where the network is unreliable its prediction is not used: predictions on
one side of an entropy threshold, calibrated on held-out maps so that the
accepted predictions meet a KS target, fall back to shooting the map (plain
IPM). The calibration also picks the side (see calibrate).
"""

import ctypes

import numpy as np

from .lib_surrogate import lib

NIN = 6
NBINS = 128
F = ctypes.POINTER(ctypes.c_float)

# high-mass power-law slope of each mass function (d N / d m ~ m^slope):
# kroupa.cuh:22-102 (above 0.5 M_sun), salpeter.cuh:22, uniform.cuh
MASS_FUNCTION_SLOPE = {"kroupa": -2.3, "salpeter": -2.35, "uniform": 0.0, "equal": 0.0}


def features(params):
    """(n, 6) raw inputs from IPM parameter dictionaries."""
    out = np.zeros((len(params), NIN))
    for i, p in enumerate(params):
        k = p["kappa_tot"]
        ks = p.get("kappa_star", k)
        out[i] = (k, p["shear"], 1.0 - ks / k if k > 0 else 0.0, MASS_FUNCTION_SLOPE[p.get("mass_function", "equal")],
                  p.get("m_lower", 0.01), p.get("m_upper", 50.0))
    return out


def normalise(h):
    h = np.asarray(h, dtype=np.float64)
    s = h.sum(axis=1, keepdims=True)
    return h / np.maximum(s, 1)


def ks_distance(p, q):
    """Kolmogorov-Smirnov distance between histograms over the same bins (rows)."""
    return np.max(np.abs(np.cumsum(normalise(p), axis=1) - np.cumsum(normalise(q), axis=1)), axis=1)


def entropy(p):
    p = np.clip(np.asarray(p, dtype=np.float64), 1e-12, 1.0)
    return -np.sum(p * np.log(p), axis=1)


def entropy_threshold(h, ks, ks_target, quantile):
    """(threshold, reject_above): the entropy threshold and the side of it that
    falls back, chosen so that the accepted predictions' KS distances have the
    given quantile within ks_target with as many accepted as possible (a tie
    goes to falling back above). (nan, True) when no non-empty set meets the
    target."""
    h = np.asarray(h, dtype=np.float64)
    ks = np.asarray(ks, dtype=np.float64)
    best = (0, np.nan, True)
    for above in (True, False):
        order = np.argsort(h if above else -h, kind="stable")
        for k in range(len(order), 0, -1):
            if np.quantile(ks[order[:k]], quantile) <= ks_target:
                if k > best[0]:
                    best = (k, float(h[order[k - 1]]), above)
                break
    return best[1], best[2]


class HistogramSurrogate:
    def __init__(self, max_batch=256, seed=1):
        self.max_batch = max_batch
        self.obj = lib.surrogate_create(max_batch, seed)
        if not self.obj:
            raise Exception("Error creating the surrogate")
        self.mean = np.zeros(NIN)
        self.std = np.ones(NIN)
        self.threshold = np.inf     # accept everything until calibrated
        self.reject_above = True
        self.losses = []

    @property
    def num_params(self):
        return lib.surrogate_num_params(self.obj)

    def params(self):
        out = np.zeros(self.num_params, dtype=np.float32)
        lib.surrogate_get_params(self.obj, out.ctypes.data_as(F))
        return out

    def set_params(self, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        if not lib.surrogate_set_params(self.obj, v.ctypes.data_as(F)):
            raise Exception("Error setting surrogate parameters")

    def _inputs(self, x):
        return np.ascontiguousarray((np.asarray(x, dtype=np.float64) - self.mean) / self.std, dtype=np.float32)

    def fit(self, x, hist, epochs=200, batch=256, lr=2e-3, wd=1e-4, seed=0, log=None):
        """Adam on the mean cross-entropy of (x, normalised hist); cosine-decayed learning rate."""
        x = np.asarray(x, dtype=np.float64)
        self.mean = x.mean(axis=0)
        self.std = np.where(x.std(axis=0) > 0, x.std(axis=0), 1.0)
        t = np.ascontiguousarray(normalise(hist), dtype=np.float32)
        xs = self._inputs(x)
        n = len(xs)
        if not lib.surrogate_load_data(self.obj, xs.ctypes.data_as(F), t.ctypes.data_as(F), n):
            raise Exception("Error loading the surrogate training set")
        batch = min(batch, self.max_batch) // 16 * 16
        rng = np.random.default_rng(seed)
        steps = epochs * (n // batch)
        step = 0
        for ep in range(epochs):
            perm = rng.permutation(n).astype(np.int32)
            tot = 0.0
            for b0 in range(0, n - batch + 1, batch):
                rows = np.ascontiguousarray(perm[b0:b0 + batch])
                rate = lr * 0.5 * (1 + np.cos(np.pi * step / max(steps, 1)))
                loss = lib.surrogate_train_step(self.obj, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), batch,
                                                rate, wd)
                if loss < 0:
                    raise Exception("Error in a surrogate training step")
                tot += loss
                step += 1
            self.losses.append(tot / max(n // batch, 1))
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log(f"[surrogate] epoch {ep}: loss {self.losses[-1]:.5f}")
        return self.losses

    def predict(self, x):
        xs = self._inputs(x)
        out = np.zeros((len(xs), NBINS), dtype=np.float32)
        if not lib.surrogate_predict(self.obj, xs.ctypes.data_as(F), len(xs), out.ctypes.data_as(F)):
            raise Exception("Error in surrogate prediction")
        return out

    def seconds_per_histogram(self, x, reps=20):
        xs = self._inputs(x)
        return lib.surrogate_time_predict(self.obj, xs.ctypes.data_as(F), len(xs), reps)

    def calibrate(self, x, hist, ks_target=0.02, quantile=0.5):
        """Entropy threshold for the fallback, set on held-out draws (x, hist).

        The predictions are ranked by entropy. For each side of a threshold
        (fall back above it, or below it) the largest accepted set whose KS
        distances to the maps' histograms have the given quantile within
        ks_target is found; the side that accepts more wins, and a tie goes to
        falling back above the threshold. On the +-50 theta_E, 256^2 maps the KS
        error of a prediction falls as its entropy rises (a peaked histogram has
        a steep CDF, so a small shift in location is a large KS distance), and
        the fallback then takes the low-entropy predictions. Returns the
        threshold; self.reject_above records the side. Accepts nothing (and
        returns nan) when no non-empty set meets the target."""
        p = self.predict(x)
        self.threshold, self.reject_above = entropy_threshold(entropy(p), ks_distance(p, hist), ks_target, quantile)
        return self.threshold

    def accept(self, p):
        """True where the prediction is used, False where the map must be shot."""
        h = entropy(p)
        if not np.isfinite(self.threshold):
            return np.zeros(len(h), dtype=bool)
        return h <= self.threshold if self.reject_above else h >= self.threshold

    def close(self):
        if self.obj:
            lib.surrogate_free(self.obj)
            self.obj = None
