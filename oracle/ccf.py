"""FP64 restatement of the critical curve and caustic finder -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of irs_amd/csrc/ccf.hip. The
product never imports it.

It follows ccf_functions.cuh and ccf.cuh of weisluke/irs with exact sums over
all stars in place of the tree and its expansions (the library's tree/FMM
agrees with the direct sums to within its alpha_error of 1e-7 Einstein radii):

* f(z, phi) = gamma + theta^2 sum m/(z - z_i)^2 - conj(d alpha_smooth/dzbar)
              - (1 - kappa - d alpha_smooth/dz) e^{-i phi}      (ccf_functions.cuh:64-80)
* f'(z)     = -2 theta^2 sum m/(z - z_i)^3 - conj(d2 alpha_smooth/dzbar2)   (116-131)
* Aberth step z <- z - f / (f' + f (2 sum 1/(z - z_i) [+ 2/z] - sum_{other roots} 1/(z - z_j)))
  with the 1/mu and step-size stopping rules of 186-240 and 309-386. The
  roots of one phase are updated simultaneously (Jacobi), as in the library.
* the phase schedule of ccf.cuh:906-1050: branch centres pi/nb + 2 pi c/nb,
  30 iterations there, then 20 per step of 2 pi/nphi outwards on both sides.
* caustics: the lens equation at the roots (ccf_functions.cuh:496-516); the
  magnification length scale 1 / (2 |caustic tangent|) (134-160).

Pinned by tests/test_ccf.py: the single-star closed form of the roots and
caustics, finite differences of the smooth-matter deflection for its first and
second derivatives, and a finite-difference construction of the magnification
length scale from the lens equation and det J.
"""

import numpy as np

from . import oracle as O


def smooth_terms(z, kappastar, rectangular, approx, taylor_smooth, corner):
    """(d alpha_smooth/dz [real], d alpha_smooth/dzbar, d2 alpha_smooth/dzbar2) at z
    (alpha_smooth.cuh:142-316). For the exact circular sheet the reference uses
    the outside forms kappa_* R^2 / zbar^2 and -2 kappa_* R^2 / zbar^3 at every
    z, also inside the circle where the sheet is linear; so does this."""
    z = np.asarray(z, dtype=complex)
    c = complex(*corner)
    zb = np.conj(z)
    if rectangular:
        box = ((-c.real <= z.real) & (z.real <= c.real) & (-c.imag <= z.imag) & (z.imag <= c.imag)).astype(float)
        if approx:
            ddz = np.full(z.shape, -kappastar)
            r = zb / c
            ph = np.angle(c)
            d1 = np.zeros(z.shape, dtype=complex)
            d2 = np.zeros(z.shape, dtype=complex)
            for i in range(taylor_smooth - 1, 1, -2):
                e = np.exp(2j * ph * i)
                d1 = (d1 + (1 - e) / i) * r * r
                d2 = (d2 + (1 - e)) * r * r
            d1 = d1 * 2 * (-1j * kappastar / np.pi) + kappastar - 4 * kappastar * ph / np.pi
            d2 = d2 / zb * 2 * (-1j * kappastar / np.pi)
        else:
            ddz = -kappastar * box
            c1, c2, c3, c4 = np.conj(c) - zb, c - zb, -c - zb, -np.conj(c) - zb
            d1 = (np.log(c1) - np.log(c2) - np.log(c3) + np.log(c4)) * (-1j * kappastar / np.pi) - kappastar * box
            d2 = (-1 / c1 + 1 / c2 + 1 / c3 - 1 / c4) * (-1j * kappastar / np.pi)
        return ddz, d1, d2
    if approx:
        return np.full(z.shape, -kappastar), np.zeros(z.shape, dtype=complex), np.zeros(z.shape, dtype=complex)
    R2 = abs(c) ** 2
    ddz = -kappastar * (abs(c) - np.abs(z) >= 0).astype(float)
    return ddz, kappastar * R2 / zb ** 2, -2 * kappastar * R2 / zb ** 3


def star_sums(z, stars):
    """sum m/(z - z_i)^2, sum m/(z - z_i)^3 and sum 1/(z - z_i) over all stars."""
    z = np.asarray(z, dtype=complex)
    s = stars[:, 0] + 1j * stars[:, 1]
    d = z[..., None] - s
    u = 1 / d
    m = stars[:, 2]
    return (m * u * u).sum(-1), (m * u * u * u).sum(-1), u.sum(-1)


class Problem:
    """A star field and its lens parameters (the fields of the CCF class)."""

    def __init__(self, stars, kappa, gamma, theta, kappastar, corner, rectangular=0, approx=1, taylor_smooth=1):
        self.stars = np.asarray(stars, dtype=np.float64).reshape(-1, 3)
        self.kappa, self.gamma, self.theta, self.kappastar = kappa, gamma, theta, kappastar
        self.corner = tuple(corner)
        self.rectangular, self.approx, self.taylor_smooth = int(rectangular), int(approx), int(taylor_smooth)

    def terms(self, z):
        S1, S2, P = star_sums(z, self.stars)
        ddz, d1, d2 = smooth_terms(z, self.kappastar, self.rectangular, self.approx, self.taylor_smooth, self.corner)
        return S1, S2, P, ddz, d1, d2

    def f0(self, z, phi):
        S1, _, _, ddz, d1, _ = self.terms(z)
        return self.gamma + self.theta ** 2 * S1 - np.conj(d1) - (1 - self.kappa - ddz) * np.exp(-1j * phi)

    def inv_mu_error(self, z, phi):
        _, _, _, ddz, _, _ = self.terms(z)
        af = np.abs(self.f0(z, phi))
        D = 1 - self.kappa - ddz
        return np.maximum(np.abs(af * (af + 2 * D)), np.abs(af * (af - 2 * D)))

    def lens(self):
        return O.make_lens(self.kappa, self.gamma, self.theta, self.kappastar, self.corner, self.rectangular,
                           self.approx, self.taylor_smooth)

    def source(self, z):
        """The lens equation w(z) by direct sums (the C oracle)."""
        z = np.asarray(z, dtype=complex)
        return O.deflect_direct(self.stars, self.lens(), z.ravel()).reshape(z.shape)

    def det_jacobian(self, z):
        S1, _, _, ddz, d1, _ = self.terms(z)
        dwdz = 1 - self.kappa - ddz
        dwdzb = self.gamma + self.theta ** 2 * np.conj(S1) - d1
        return dwdz ** 2 - np.abs(dwdzb) ** 2

    def mu_length_scale(self, z):
        S1, S2, _, ddz, d1, d2 = self.terms(z)
        dwdz = 1 - self.kappa - ddz
        dwdzb = self.gamma + self.theta ** 2 * np.conj(S1) - d1
        d2w = -np.conj(2 * self.theta ** 2 * S2) - d2
        cct = -2j * np.conj(dwdzb) * d2w
        ct = dwdz * cct + dwdzb * np.conj(cct)
        return 1 / (2 * np.abs(ct))

    def step(self, z, phi, done):
        """One simultaneous Aberth update of the roots z (one phase); returns (z, done)."""
        S1, S2, P, ddz, d1, d2 = self.terms(z)
        f0 = self.gamma + self.theta ** 2 * S1 - np.conj(d1) - (1 - self.kappa - ddz) * np.exp(-1j * phi)
        af = np.abs(f0)
        D = 1 - self.kappa - ddz
        small = np.maximum(np.abs(af * (af + 2 * D)), np.abs(af * (af - 2 * D))) < 1e-9
        f1 = -2 * self.theta ** 2 * S2 - np.conj(d2)
        psum = 2 * P
        if not self.rectangular and not self.approx:
            psum = psum + 2 / z
        diff = z[:, None] - z[None, :]
        np.fill_diagonal(diff, 1)
        inv = 1 / diff
        np.fill_diagonal(inv, 0)
        rsum = inv.sum(1)
        zn = z - f0 / (f1 + f0 * (psum - rsum))
        upd = ~done & ~small
        new_done = done | small | (upd & (np.abs(zn - z) / self.theta < 1e-9))
        return np.where(upd, zn, z), new_done


def num_roots(num_stars, rectangular, approx, taylor_smooth):
    """ccf.cuh:463-483."""
    if rectangular:
        return 2 * num_stars + taylor_smooth - 1 if approx else 2 * num_stars
    return 2 * num_stars if approx else 2 * num_stars + 2


def initial_roots(stars, nroots, corner):
    """ccf.cuh:690-703: each star +-1, the rest evenly on the circle |z| = |corner|."""
    s = stars[:, 0] + 1j * stars[:, 1]
    extra = nroots - 2 * len(s)
    ring = abs(complex(*corner)) * np.exp(2j * np.pi * np.arange(extra) / max(extra, 1))
    return np.concatenate([s + 1, s - 1, ring[:extra]])


def solve(prob, nphi, nb, init_iters=30, step_iters=20, init=None):
    """Roots of every phase, root-major: (nroots, nphi + nb) complex, plus the
    row phases and the largest error bound on 1/mu."""
    n = num_roots(len(prob.stars), prob.rectangular, prob.approx, prob.taylor_smooth)
    z0 = initial_roots(prob.stars, n, prob.corner) if init is None else np.asarray(init, dtype=complex)
    J = nphi // (2 * nb)
    rows = np.zeros((nphi + nb, n), dtype=complex)
    phis = np.zeros(nphi + nb)
    for c in range(nb):
        centre = c * (2 * J + 1) + J
        phi0 = np.pi / nb + c * 2 * np.pi / nb
        for s in range(-J, J + 1):
            phis[centre + s] = phi0 + s * 2 * np.pi / nphi
        z, done = z0.copy(), np.zeros(n, dtype=bool)
        for _ in range(init_iters):
            z, done = prob.step(z, phi0, done)
        rows[centre] = z
        for sgn in (-1, 1):
            z = rows[centre].copy()
            for j in range(1, J + 1):
                done = np.zeros(n, dtype=bool)
                phi = phi0 + sgn * j * 2 * np.pi / nphi
                for _ in range(step_iters):
                    z, done = prob.step(z, phi, done)
                rows[centre + sgn * j] = z
    err = np.array([prob.inv_mu_error(rows[r], phis[r]) for r in range(nphi + nb)])
    return rows.T.copy(), phis, float(err.max())
