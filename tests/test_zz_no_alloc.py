"""ipm_deflect_noalloc (include/lib_ipm.h) against ipm_deflect through the C
ABI: the same kernel on the same points, so the deflections and
magnifications must be identical, on calls in a row of decreasing size (the
staging array keeps the last call's contents), at the largest size it takes,
and in both libraries; sizes beyond that and calls before build() are
refused on the host."""

import ctypes
import sys

import numpy as np
import pytest

from common import C1_SMALL, from_starfile, random_points, write_starfile

pytestmark = pytest.mark.gpu

P = ctypes.POINTER(ctypes.c_double)
MAX_POINTS = 65536


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def call(lib, obj, name, z, want_mu):
    zz = np.ascontiguousarray(np.column_stack([z.real, z.imag]))
    w = np.zeros((len(z), 2))
    mu = np.zeros(len(z)) if want_mu else None
    ok = getattr(lib, name)(obj, zz.ctypes.data_as(P), len(z), w.ctypes.data_as(P),
                            mu.ctypes.data_as(P) if want_mu else None)
    return ok, w, mu


@pytest.fixture(scope="module")
def starfile(oracle, tmp_path_factory):
    path = tmp_path_factory.mktemp("noalloc") / "stars.bin"
    write_starfile(oracle, path, C1_SMALL, seed=5)
    return path


@pytest.mark.parametrize("is_double", [True, False])
def test_noalloc_deflection_equals_ipm_deflect(starfile, is_double):
    from irs_amd.microlensing.IPM.ipm import IPM
    m = IPM(**from_starfile(C1_SMALL, starfile), is_double=is_double, expansion_order=12)
    m.prepare()
    # before build() there is no tree: refused on the host
    ok, _, _ = call(m.lib, m.obj, "ipm_deflect_noalloc", np.zeros(4, dtype=complex), False)
    assert not ok
    m.build()
    for n, seed in ((5000, 1), (1200, 2), (MAX_POINTS, 3), (1, 4)):
        z = random_points(m, n, seed=seed)
        for want_mu in ((True, False) if is_double else (False,)):
            say(f"[gpu] noalloc deflection: is_double={is_double} n={n} mu={want_mu}")
            ok_a, w_a, mu_a = call(m.lib, m.obj, "ipm_deflect", z, want_mu)
            ok_b, w_b, mu_b = call(m.lib, m.obj, "ipm_deflect_noalloc", z, want_mu)
            assert ok_a and ok_b
            np.testing.assert_array_equal(w_b, w_a)
            if want_mu:
                np.testing.assert_array_equal(mu_b, mu_a)
            assert np.isfinite(w_b).all()
    # more points than the staging array holds: refused before any device work
    ok, _, _ = call(m.lib, m.obj, "ipm_deflect_noalloc", np.zeros(MAX_POINTS + 1, dtype=complex), False)
    assert not ok
    ok, w, _ = call(m.lib, m.obj, "ipm_deflect_noalloc", np.zeros(0, dtype=complex), False)
    assert ok and w.size == 0
    m.close()
