"""FP64 CPU oracle for the overlap-maximisation hot path (test infrastructure).

This module is TEST INFRASTRUCTURE ONLY.  It restates, in plain numpy/scipy
FP64, the algorithm of the reference's pure-Python classes so that the HIP
implementation in ``fastoverlap_amd`` can be checked against it.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it, and only as the checker (or as the timed CPU baseline),
never as the thing measured or shipped.

Every function cites the reference file:line it restates.  Paths are relative
to the reference repository root (``fastoverlap/...``).

Pinning (what checks each function against something other than itself) is
listed in ``PINNED`` at the bottom of this file and exercised by
``tests/test_oracle.py``.  Every function there is pinned.  Two of them are
pinned by what the mathematics fixes rather than by a reference output:

* ``max_overlap`` -- the end point of scipy's L-BFGS-B is scipy-version
  specific, so it is pinned by the optimum itself: an exact rotated copy must
  be mapped back onto the true rotation, the end point of a noisy pair must be
  a stationary local minimum that an independent (Nelder-Mead) search agrees
  with;
* ``PeriodicOracle.hungarian`` -- implements the *intended* LAP semantics
  (rows = X1, ``X2[perm]`` ~ ``X1``); the reference's own munkres fallback
  applies the inverse permutation (``fastoverlap/periodicAlignment.py:94-107``)
  and its published BLJ256 answer was produced with pele, absent here.  It is
  pinned by brute force over every permutation, by exact shifted copies and
  by that BLJ256 answer (distance 1.559).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from numpy import pi
from scipy.optimize import linear_sum_assignment, minimize
from scipy.special import eval_jacobi, gammaln, ive, sph_harm_y

# ---------------------------------------------------------------------------
# SO(3) grid, quadrature weights and Wigner-d tables   (fastoverlap/soft.py)
# ---------------------------------------------------------------------------


def soft_grid(bw: int):
    """Grid angles (alpha, beta, gamma) of a bandwidth-``bw`` SO(3) grid.

    Restates ``fastoverlap/soft.py:48-56``: alpha_j = pi j / bw,
    beta_k = pi (2k+1) / (4 bw), gamma_j = pi j / bw for j, k in [0, 2bw).
    """
    j = np.arange(2 * bw)
    return pi / bw * j, pi / 4 / bw * (2 * j + 1), pi / bw * j


def makeweights(bw: int) -> np.ndarray:
    """Quadrature weights of the beta grid (``fastoverlap/soft.py:63-71``)."""
    j = np.arange(0, 2 * bw).astype(float)[:, None]
    k = np.arange(0, bw).astype(float)[None, :]
    fudge = pi / 4 / bw
    return (2 / (2 * k + 1) * np.sin((2 * j + 1) * (2 * k + 1) * fudge)
            * np.sin((2 * j + 1) * fudge) / bw).sum(1)


def lmm_triples(L: int):
    """(l, m1, m2) for l <= L, m1, m2 in [-l, l] in the reference's order
    (``fastoverlap/soft.py:57-59``; ``fastoverlap/sphericalAlignment.py:52-55``)."""
    out = [(l, m1, m2) for l in range(L + 1)
           for m1 in range(-l, l + 1) for m2 in range(-l, l + 1)]
    Js, m1s, m2s = (np.array(v) for v in zip(*out))
    return Js, m1s, m2s


def _log_factorial(x):
    return gammaln(np.asarray(x, float) + 1.0)


def wigner_d_jacobi(L: int, beta, normalised: bool = True):
    """Wigner small-d via the Jacobi-polynomial closed form.

    Restates ``fastoverlap/soft.py:73-96`` (normalised, times sqrt((2l+1)/2))
    and ``fastoverlap/sphericalAlignment.py:67-85`` (unnormalised).  The
    factorial ratio is evaluated through log-gamma, which equals the
    reference's ``sqrt(factorial(...)/...)`` where that does not overflow.

    Returns ``d[l, m1 % (2L+1), m2 % (2L+1), k]`` for ``beta`` of shape (nb,).
    """
    beta = np.atleast_1d(np.asarray(beta, float))
    Js, m1s, m2s = lmm_triples(L)
    mu = np.abs(m1s - m2s)
    nu = np.abs(m1s + m2s)
    s = Js - (mu + nu) // 2
    xi = np.ones(len(s))
    neg = m2s < m1s
    xi[neg] = (-1.0) ** (m1s - m2s)[neg]
    logf = 0.5 * (_log_factorial(s) + _log_factorial(s + mu + nu)
                  - _log_factorial(s + mu) - _log_factorial(s + nu))
    factor = np.exp(logf) * xi
    cosb = np.cos(beta)
    jac = eval_jacobi(s[:, None], mu[:, None], nu[:, None], cosb[None, :])
    val = (factor[:, None] * jac * np.sin(beta / 2)[None, :] ** mu[:, None]
           * np.cos(beta / 2)[None, :] ** nu[:, None])
    if normalised:
        val *= np.sqrt((2.0 * Js[:, None] + 1.0) / 2.0)
    n = 2 * L + 1
    out = np.zeros((L + 1, n, n, len(beta)))
    out[Js, m1s % n, m2s % n, :] = val
    return out


def wigner_d_recurrence(L: int, beta, normalised: bool = True):
    """FP64 three-term recurrence in l for the Wigner small-d.

    Method of Kostelec & Rockmore (2008), the one the reference's Fortran
    uses (``f90/DSOFT.f90:81-195``): seed d^{l0}_{m1 m2} at
    l0 = max(|m1|, |m2|) in closed form, then

        d^{l+1} = B_l (cos b - C_l) d^l - A_l d^{l-1}

    with the normalised-d coefficients of ``_recurrence_terms`` (derived from
    the paper's unnormalised relation).  Used by the oracle at high l where
    the Jacobi closed form loses accuracy.  Same output layout as
    ``wigner_d_jacobi``.
    """
    beta = np.atleast_1d(np.asarray(beta, float))
    n = 2 * L + 1
    out = np.zeros((L + 1, n, n, len(beta)))
    c2 = np.cos(beta / 2)
    s2 = np.sin(beta / 2)
    cb = np.cos(beta)
    for m1 in range(-L, L + 1):
        for m2 in range(-L, L + 1):
            l0 = max(abs(m1), abs(m2))
            seed = _wigner_seed(l0, m1, m2, c2, s2)  # normalised d^{l0}
            prev = np.zeros_like(seed)
            cur = seed
            out[l0, m1 % n, m2 % n] = cur
            for l in range(l0, L):
                A, B, C = _recurrence_terms(l, m1, m2)
                nxt = B * (cb - C) * cur - A * prev
                prev, cur = cur, nxt
                out[l + 1, m1 % n, m2 % n] = cur
    if not normalised:
        out *= np.sqrt(2.0 / (2.0 * np.arange(L + 1) + 1.0))[:, None, None, None]
    return out


def _recurrence_terms(l: int, m1: int, m2: int):
    """(A, B, C) of dn^{l+1} = B (cos b - C) dn^l - A dn^{l-1} for the
    normalised dn^l = sqrt((2l+1)/2) d^l (Kostelec & Rockmore 2008).

    From the unnormalised three-term relation with the edge factor
    e_j = sqrt((j^2 - m1^2)(j^2 - m2^2)):
    e_{j+1} d^{j+1} = (j+1)(2j+1)(cos b - m1 m2/(j(j+1))) d^j - (j+1)/j e_j d^{j-1}.
    """
    j = float(l)
    e_next = math.sqrt(((j + 1) ** 2 - m1 * m1) * ((j + 1) ** 2 - m2 * m2))
    B = (j + 1) * math.sqrt((2 * j + 1) * (2 * j + 3)) / e_next
    if l == 0:
        return 0.0, B, 0.0
    e_cur = math.sqrt((j * j - m1 * m1) * (j * j - m2 * m2))
    A = (j + 1) / j * math.sqrt((2 * j + 3) / (2 * j - 1)) * e_cur / e_next
    return A, B, m1 * m2 / (j * (j + 1))


def _wigner_seed(l0: int, m1: int, m2: int, c2, s2):
    """Normalised d^{l0}_{m1 m2}(beta) at l0 = max(|m1|,|m2|) in closed form.

    For the edge element the Jacobi polynomial has degree 0, so the closed
    form of ``fastoverlap/soft.py:82-95`` reduces to a single monomial in
    cos(b/2), sin(b/2); this is the seed used by ``DSOFT.f90:150-162``.
    """
    mu = abs(m1 - m2)
    nu = abs(m1 + m2)
    s = l0 - (mu + nu) // 2  # == 0 at the edge
    xi = (-1.0) ** (m1 - m2) if m2 < m1 else 1.0
    logf = 0.5 * (math.lgamma(s + 1) + math.lgamma(s + mu + nu + 1)
                  - math.lgamma(s + mu + 1) - math.lgamma(s + nu + 1))
    return (math.sqrt((2 * l0 + 1) / 2.0) * xi * math.exp(logf)
            * s2 ** mu * c2 ** nu)


class SOFTOracle:
    """Restatement of ``SOFT`` (``fastoverlap/soft.py:40-130``).

    ``Ds[l, m1, m2, k]`` is the dense normalised table (``soft.py:73-96``);
    ``method='jacobi'`` uses the reference's closed form, ``'recurrence'`` the
    FP64 recurrence (needed where the closed form loses accuracy).
    """

    def __init__(self, bw: int, method: str = "jacobi"):
        self.bw = bw
        self.n = 2 * bw
        self.Jmax = bw - 1
        self.weights = makeweights(bw)
        self.a, self.b, self.y = soft_grid(bw)
        if method == "jacobi":
            self.Ds = wigner_d_jacobi(self.Jmax, self.b)
        else:
            self.Ds = wigner_d_recurrence(self.Jmax, self.b)
        self.indFactor = np.array([2 * pi / self.n, pi / self.n, 2 * pi / self.n])

    def iSOFT(self, flmm):
        """Inverse SOFT (``fastoverlap/soft.py:115-125``)."""
        Jmax, bw = self.Jmax, self.bw
        assert flmm.shape == (bw, bw * 2 - 1, bw * 2 - 1)
        S2out = np.zeros((2 * bw, 2 * bw, 2 * bw), np.complex128)
        for m1 in range(-Jmax, Jmax + 1):
            for m2 in range(-Jmax, Jmax + 1):
                l = max(abs(m1), abs(m2))
                S2out[m1, :, m2] = self.Ds[l:, m1, m2].T.dot(flmm[l:, m1, m2])
        S1out = np.fft.ifft(S2out, axis=2)
        return np.fft.ifft(S1out, axis=0) * (2 * bw) ** 2

    def SOFT(self, data):
        """Forward SOFT (``fastoverlap/soft.py:98-113``)."""
        Jmax, bw = self.Jmax, self.bw
        data = np.asanyarray(data)
        assert all(n == self.n for n in data.shape)
        S1 = np.fft.fft(data, axis=0)
        S2 = np.fft.fft(S1, axis=2) * (2.0 * bw) ** -2
        flmm = np.zeros((bw, bw * 2 - 1, bw * 2 - 1), np.complex128)
        for m1 in range(-Jmax, Jmax + 1):
            for m2 in range(-Jmax, Jmax + 1):
                l = max(abs(m1), abs(m2))
                flmm[l:, m1, m2] = self.Ds[l:, m1, m2].dot(self.weights * S2[m1, :, m2])
        return flmm

    def indtoEuler(self, ind):
        """Grid index -> Euler angles (``fastoverlap/soft.py:127-130``)."""
        R = self.indFactor * np.atleast_2d(np.asarray(ind, float))
        R[:, 1] += 0.5 * pi / self.n
        return R.squeeze()


def pad_coeffs(I, bw: int):
    """Zero-pad C^l (L+1, 2L+1, 2L+1) to a SOFT(bw) input (bw, 2bw-1, 2bw-1).

    Needed when the grid is larger than 2(L+1) (batched configurations); the
    reference's ``SOFT`` accepts any ``bw`` (``fastoverlap/soft.py:115-118``).
    """
    L = I.shape[0] - 1
    n = 2 * bw - 1
    out = np.zeros((bw, n, n), np.complex128)
    for l in range(L + 1):
        for m1 in range(-l, l + 1):
            for m2 in range(-l, l + 1):
                out[l, m1 % n, m2 % n] = I[l, m1, m2]
    return out


# ---------------------------------------------------------------------------
# Coefficients C^l_{m1 m2}           (fastoverlap/sphericalAlignment.py)
# ---------------------------------------------------------------------------


def calc_theta_phi_r(pos):
    """``fastoverlap/utils.py:439-445``."""
    pos = np.atleast_2d(pos)
    X, Y, Z = pos.T
    R = np.linalg.norm(pos, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        theta = np.arccos(Z / R)
    return theta, np.arctan2(Y, X), R


def sph_harm_table(L: int, theta, phi):
    """Y[l, m % (2L+1), j] (``fastoverlap/sphericalAlignment.py:57-65``).

    scipy's orthonormal, Condon-Shortley-phased Y_l^m -- the same library
    routine family the reference calls (``sph_harm(m, l, phi, theta)`` is
    ``sph_harm_y(l, m, theta, phi)``).
    """
    theta = np.atleast_1d(theta)
    phi = np.atleast_1d(phi)
    Jinds = np.array([j for j in range(L + 1) for _ in range(-j, j + 1)])
    Minds = np.array([m for j in range(L + 1) for m in range(-j, j + 1)])
    Y = np.zeros((L + 1, 2 * L + 1) + theta.shape, np.complex128)
    Y[Jinds, Minds] = sph_harm_y(Jinds[:, None], Minds[:, None],
                                 theta[None, :], phi[None, :])
    return Y


def radial_kernel(L: int, r1, r2, scale: float):
    """il[l, j, k] = i_l(x) exp(-(r1^2+r2^2)/4s^2), x = r1 r2 / 2s^2.

    Restates ``fastoverlap/sphericalAlignment.py:269-272``.  The reference
    evaluates ``iv(l+1/2, x) sqrt(pi/2x) exp(-(r1^2+r2^2)/4s^2)``, which
    overflows for x > ~700; this uses the identical quantity in the
    exponentially scaled form ``ive(l+1/2, x) sqrt(pi/2x) exp(-(r1-r2)^2/4s^2)``
    (since (r1^2+r2^2)/4s^2 = (r1-r2)^2/4s^2 + x).  At x = 0 the reference
    yields NaN (0 * inf); here the i_0(0) = 1, i_l(0) = 0 limit is used.
    """
    x = r1[:, None] * r2[None, :] / scale ** 2 / 2
    ls = np.arange(L + 1)[:, None, None]
    with np.errstate(invalid="ignore", divide="ignore"):
        il = ive(ls + 0.5, x[None]) * np.sqrt(pi / 2 / x[None])
    zero = x == 0
    if np.any(zero):
        il[:, zero] = 0.0
        il[0, zero] = 1.0
    il *= np.exp(-(r1[:, None] - r2[None, :]) ** 2 / 4 / scale ** 2)[None]
    return il


def so3_coeffs(pos1, pos2, L: int, scale: float, fast: bool = False):
    """C^l_{m1 m2} for one species (``fastoverlap/sphericalAlignment.py:260-273``).

    I[l,m1,m2] = 4 pi^{5/2} s^3 sum_{j,k} il[l,j,k] Y1[l,m1,j] conj(Y2[l,m2,k]).
    ``fast=False`` evaluates the reference's naive ``einsum``; ``fast=True``
    the identical sum as the GEMM chain (Y1 . il_l) . Y2^H (for big tests).
    """
    pos1 = np.atleast_2d(pos1)
    pos2 = np.atleast_2d(pos2)
    assert pos1.shape == pos2.shape
    t1, p1, r1 = calc_theta_phi_r(pos1)
    t2, p2, r2 = calc_theta_phi_r(pos2)
    # atoms exactly at the centre: any direction gives the same (l=0) limit
    t1 = np.where(r1 == 0, 0.0, t1)
    t2 = np.where(r2 == 0, 0.0, t2)
    Y1 = sph_harm_table(L, t1, p1)
    Y2 = sph_harm_table(L, t2, p2).conj()
    il = radial_kernel(L, r1, r2, scale)
    if fast:
        out = np.empty((L + 1, 2 * L + 1, 2 * L + 1), np.complex128)
        for l in range(L + 1):
            out[l] = (Y1[l] @ il[l]) @ Y2[l].T
    else:
        out = np.einsum("ijk,ilj,imk->ilm", il, Y1, Y2)
    return out * 4 * np.pi ** 2.5 * scale ** 3


def so3_coeffs_species(X1, X2, perm, L, scale, fast=False):
    """sum over permutation groups (``fastoverlap/sphericalAlignment.py:175``)."""
    return sum(so3_coeffs(X1[p], X2[p], L, scale, fast) for p in perm)


def brute_overlap(pos1, pos2, scale):
    """Direct Gaussian overlap (``fastoverlap/utils.py:402-406``)."""
    d = np.linalg.norm(pos1[:, None, :] - pos2[None, :, :], axis=2)
    return np.exp(-d ** 2 / 4 / scale ** 2).sum() * (pi * scale ** 2) ** 1.5


# ---------------------------------------------------------------------------
# Peak search and rotation refinement
# ---------------------------------------------------------------------------


def find_max(a):
    """Interpolated argmax (``fastoverlap/utils.py:319-338``).

    numpy's argmax (lexicographic on complex: real, then imaginary; first
    occurrence in C order), then a per-axis parabolic offset on |a| with
    periodic wrap.  Returns (integer index, fractional index).
    """
    a = np.asanyarray(a)
    shape = a.shape
    dim = len(shape)
    ind = np.unravel_index(a.argmax(), shape)
    i1 = tuple(tuple((i0 + 1) % shape[i] if i == j else i0 for j in range(dim))
               for i, i0 in enumerate(ind))
    i2 = tuple(tuple(i0 for j in range(dim)) for i, i0 in enumerate(ind))
    i3 = tuple(tuple(i0 - 1 if i == j else i0 for j in range(dim))
               for i, i0 in enumerate(ind))
    y1 = np.abs(a[i1])
    y2 = np.abs(a[i2])
    y3 = np.abs(a[i3])
    d = (y3 - y1) / (2 * (2 * y2 - y1 - y3))
    return np.array(ind), np.array(ind) - d


def wigner_D_rot(L: int, rot):
    """D^l_{m1m2}(a,b,y) and its gradient (``fastoverlap/sphericalAlignment.py:67-96``).

    D = exp(-i m1 a) d^l_{m1m2}(b) exp(-i m2 y) with the *unnormalised* d;
    the beta derivative uses the Jacobi derivative (``utils.py:398-400``).
    ``_wigner_d_and_derivs`` gives the same derivative from the
    differentiated recurrence (checked against it and finite differences).
    """
    return wigner_D_ref(L, rot)


def _wigner_d_and_derivs(L: int, b: float):
    """Unnormalised d, d', d'' at a single beta by the recurrence, exact derivs."""
    n = 2 * L + 1
    d = np.zeros((L + 1, n, n))
    d1 = np.zeros_like(d)
    d2 = np.zeros_like(d)
    cb, sb = math.cos(b), math.sin(b)
    c2, s2 = math.cos(b / 2), math.sin(b / 2)
    for m1 in range(-L, L + 1):
        for m2 in range(-L, L + 1):
            l0 = max(abs(m1), abs(m2))
            mu, nu = abs(m1 - m2), abs(m1 + m2)
            xi = (-1.0) ** (m1 - m2) if m2 < m1 else 1.0
            s = l0 - (mu + nu) // 2
            logf = 0.5 * (math.lgamma(s + 1) + math.lgamma(s + mu + nu + 1)
                          - math.lgamma(s + mu + 1) - math.lgamma(s + nu + 1))
            f = xi * math.exp(logf)
            # g(b) = s2^mu c2^nu ; derivatives w.r.t. b
            g = s2 ** mu * c2 ** nu
            g1 = 0.5 * (mu * (s2 ** (mu - 1) if mu else 0.0) * c2 ** (nu + 1)
                        - nu * s2 ** (mu + 1) * (c2 ** (nu - 1) if nu else 0.0))
            # second derivative via product rule on the two terms
            def term(p, q):  # d/db of s2^p c2^q
                a1 = 0.5 * p * (s2 ** (p - 1) if p else 0.0) * c2 ** (q + 1)
                a2 = -0.5 * q * s2 ** (p + 1) * (c2 ** (q - 1) if q else 0.0)
                return a1 + a2
            g2 = 0.5 * (mu * term(mu - 1, nu + 1) if mu else 0.0) \
                - 0.5 * (nu * term(mu + 1, nu - 1) if nu else 0.0)
            # normalised recurrence, converted to unnormalised at the end
            nl0 = math.sqrt((2 * l0 + 1) / 2.0)
            cur = np.array([f * g, f * g1, f * g2]) * nl0
            prev = np.zeros(3)
            for l in range(l0, L + 1):
                un = math.sqrt(2.0 / (2 * l + 1))
                d[l, m1 % n, m2 % n] = cur[0] * un
                d1[l, m1 % n, m2 % n] = cur[1] * un
                d2[l, m1 % n, m2 % n] = cur[2] * un
                if l == L:
                    break
                A, B, C = _recurrence_terms(l, m1, m2)
                nxt = np.empty(3)
                nxt[0] = B * (cb - C) * cur[0] - A * prev[0]
                nxt[1] = B * (-sb * cur[0] + (cb - C) * cur[1]) - A * prev[1]
                nxt[2] = B * (-cb * cur[0] - 2 * sb * cur[1] + (cb - C) * cur[2]) \
                    - A * prev[2]
                prev, cur = cur, nxt
    return d, d1, d2


_JAC_CACHE = {}


def _jacobi_terms(L):
    """Per-(l,m1,m2) constants of the closed form (sphericalAlignment.py:74-83)."""
    if L not in _JAC_CACHE:
        Js, m1s, m2s = lmm_triples(L)
        mu = np.abs(m1s - m2s)
        nu = np.abs(m1s + m2s)
        s = Js - (mu + nu) // 2
        xi = np.ones(len(s))
        neg = m2s < m1s
        xi[neg] = (-1.0) ** (m1s - m2s)[neg]
        logf = 0.5 * (_log_factorial(s) + _log_factorial(s + mu + nu)
                      - _log_factorial(s + mu) - _log_factorial(s + nu))
        # d/dx P_s^(mu,nu)(x) = (s+mu+nu+1)/2 P_{s-1}^(mu+1,nu+1)(x)  (utils.py:398-400)
        gfact = (s + mu + nu + 1) / 2.0
        _JAC_CACHE[L] = (Js, m1s, m2s, mu, nu, s, np.exp(logf) * xi, gfact)
    return _JAC_CACHE[L]


def wigner_D_ref(L, rot):
    """D and dD/d(a,b,y) by the reference's own formulas
    (``fastoverlap/sphericalAlignment.py:67-96``), vectorised."""
    a, b, y = rot
    Js, m1s, m2s, mu, nu, s, factor, gfact = _jacobi_terms(L)
    sb2, cb2, cb, sb = math.sin(b / 2), math.cos(b / 2), math.cos(b), math.sin(b)
    jac = eval_jacobi(s, mu, nu, cb)
    gjac = np.where(s > 0, eval_jacobi(np.maximum(s - 1, 0), mu + 1, nu + 1, cb) * gfact, 0.0)
    gjac = gjac * -sb
    with np.errstate(divide="ignore", invalid="ignore"):
        p_mu1 = np.where(mu > 0, sb2 ** (mu - 1.0), 0.0)
        p_nu1 = np.where(nu > 0, cb2 ** (nu - 1.0), 0.0)
    d = factor * jac * sb2 ** mu * cb2 ** nu
    dd = (factor * gjac * sb2 ** mu * cb2 ** nu
          + factor * jac * p_mu1 * cb2 ** (nu + 1) * mu / 2
          - factor * jac * sb2 ** (mu + 1) * p_nu1 * nu / 2)
    n = 2 * L + 1
    D = np.zeros((L + 1, n, n), np.complex128)
    G = np.zeros((3, L + 1, n, n), np.complex128)
    ph = np.exp(-1j * m1s * a) * np.exp(-1j * m2s * y)
    D[Js, m1s % n, m2s % n] = ph * d
    G[0, Js, m1s % n, m2s % n] = -1j * m1s * ph * d
    G[1, Js, m1s % n, m2s % n] = ph * dd
    G[2, Js, m1s % n, m2s % n] = -1j * m2s * ph * d
    return D, G


def energy_gradient(rot, Iconj, L):
    """f(R) = -Re sum Iconj * D(R) and its gradient
    (``fastoverlap/sphericalAlignment.py:98-100``)."""
    D, G = wigner_D_ref(L, rot)
    return -(Iconj * D).real.sum(), -(Iconj[None] * G).real.sum((1, 2, 3))


def energy_gradient_recurrence(rot, Iconj, L):
    """Same quantity with d, d' from the differentiated recurrence (cross-check)."""
    a, b, y = rot
    n = 2 * L + 1
    d, d1, _ = _wigner_d_and_derivs(L, b)
    ms = np.arange(n)
    ms = np.where(ms > L, ms - n, ms)
    ea = np.exp(-1j * ms * a)[None, :, None]
    ey = np.exp(-1j * ms * y)[None, None, :]
    D = ea * d * ey
    Db = ea * d1 * ey
    f = -(Iconj * D).real.sum()
    g = np.array([-(Iconj * (-1j) * ms[None, :, None] * D).real.sum(),
                  -(Iconj * Db).real.sum(),
                  -(Iconj * (-1j) * ms[None, None, :] * D).real.sum()])
    return f, g


def max_overlap(R0, Iconj, L):
    """scipy L-BFGS-B from the grid angle
    (``fastoverlap/sphericalAlignment.py:110-113``).  Its iterates are
    scipy-version specific; the end point is pinned by the optimum itself
    (exact copies, stationarity, an independent search)."""
    res = minimize(energy_gradient, np.asarray(R0, float), jac=True,
                   args=(Iconj, L), method="L-BFGS-B")
    return res.x, res


def euler_matrix(a, b, y):
    """``fastoverlap/utils.py:447-460``: My . Mb . Ma."""
    sa, ca = math.sin(a), math.cos(a)
    sb, cb = math.sin(b), math.cos(b)
    sy, cy = math.sin(y), math.cos(y)
    Ma = np.array(((ca, -sa, 0), (sa, ca, 0), (0, 0, 1)))
    Mb = np.array(((cb, 0, -sb), (0, 1, 0), (sb, 0, cb)))
    My = np.array(((cy, -sy, 0), (sy, cy, 0), (0, 0, 1)))
    return My.dot(Mb).dot(Ma)


# ---------------------------------------------------------------------------
# Linear assignment and Kabsch (fastoverlap/utils.py)
# ---------------------------------------------------------------------------


def lap(cost):
    """Exact LAP (``fastoverlap/utils.py:34-46``): perm[i] = column of row i.
    Any exact solver agrees with the reference's munkres when the optimum is
    unique; scipy's shortest-augmenting-path solver is used here."""
    r, c = linear_sum_assignment(cost)
    perm = np.empty(len(r), int)
    perm[r] = c
    return perm


def find_best_permutation(X1, X2, permlist, cost_fn=None):
    """Per-group LAP assembled into one permutation
    (``fastoverlap/utils.py:82-167``).  Returns (dist, perm) with
    ``X2[perm]`` matched to ``X1``; dist = |X1 - X2[perm]| over the groups."""
    if cost_fn is None:
        cost_fn = lambda A, B: ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1)
    X1 = X1.reshape(-1, 3)
    X2 = X2.reshape(-1, 3)
    newperm = np.arange(len(X1))
    for atomlist in permlist:
        atomlist = np.asarray(atomlist)
        if len(atomlist) == 0:
            continue
        p = lap(cost_fn(X1[atomlist], X2[atomlist]))
        newperm[atomlist] = atomlist[p]
    return np.linalg.norm(X1 - X2[newperm]), newperm


def kabsch_quaternion(x1, x2):
    """Kearsley quaternion fit (``fastoverlap/utils.py:169-275``).
    Returns (dist, MR) with dist = sqrt(min eigenvalue of QMAT)."""
    x1 = x1.reshape(-1, 3) - x1.reshape(-1, 3).mean(0)
    x2 = x2.reshape(-1, 3) - x2.reshape(-1, 3).mean(0)
    xm = x1 - x2
    xp = x1 + x2
    XM, YM, ZM = xm.T
    XP, YP, ZP = xp.T
    Q = np.zeros((4, 4))
    Q[0, 0] = (XM ** 2 + YM ** 2 + ZM ** 2).sum()
    Q[0, 1] = (-YP * ZM + YM * ZP).sum()
    Q[0, 2] = (-XM * ZP + XP * ZM).sum()
    Q[0, 3] = (-XP * YM + XM * YP).sum()
    Q[1, 1] = (YP ** 2 + ZP ** 2 + XM ** 2).sum()
    Q[1, 2] = (XM * YM - XP * YP).sum()
    Q[1, 3] = (XM * ZM - XP * ZP).sum()
    Q[2, 2] = (XP ** 2 + ZP ** 2 + YM ** 2).sum()
    Q[2, 3] = (YM * ZM - YP * ZP).sum()
    Q[3, 3] = (XP ** 2 + YP ** 2 + ZM ** 2).sum()
    Q = np.triu(Q) + np.triu(Q, 1).T
    w, v = np.linalg.eigh(Q)
    eig = w[0]
    q = v[:, 0]
    if eig < 0:
        eig = 0.0 if abs(eig) < 1e-6 else -eig
    return math.sqrt(eig), q2mx(q)


def q2mx(qin):
    """Quaternion -> rotation matrix (``fastoverlap/utils.py:255-275``)."""
    Q = qin / np.linalg.norm(qin)
    R = np.zeros((3, 3))
    R[0, 0] = 2.0 * (0.5 - Q[2] * Q[2] - Q[3] * Q[3])
    R[1, 1] = 2.0 * (0.5 - Q[1] * Q[1] - Q[3] * Q[3])
    R[2, 2] = 2.0 * (0.5 - Q[1] * Q[1] - Q[2] * Q[2])
    R[0, 1] = 2.0 * (Q[1] * Q[2] - Q[0] * Q[3])
    R[1, 0] = 2.0 * (Q[1] * Q[2] + Q[0] * Q[3])
    R[0, 2] = 2.0 * (Q[1] * Q[3] + Q[0] * Q[2])
    R[2, 0] = 2.0 * (Q[1] * Q[3] - Q[0] * Q[2])
    R[1, 2] = 2.0 * (Q[2] * Q[3] - Q[0] * Q[1])
    R[2, 1] = 2.0 * (Q[2] * Q[3] + Q[0] * Q[1])
    return R


def average_separation(pos):
    """Mean nearest-neighbour distance (``fastoverlap/sphericalAlignment.py:35-41``)."""
    pos = np.asanyarray(pos).reshape(-1, 3)
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    np.fill_diagonal(d, np.inf)
    return d.min(1).sum() / len(pos)


# ---------------------------------------------------------------------------
# Spherical alignment driver      (fastoverlap/sphericalAlignment.py:129-273)
# ---------------------------------------------------------------------------


@dataclass
class SphericalResult:
    dist: float
    X1: np.ndarray
    X2: np.ndarray
    perm: np.ndarray
    inverted: bool
    R: np.ndarray
    ind: list = field(default_factory=list)   # integer argmax per orientation
    fun: list = field(default_factory=list)   # objective per orientation
    grids: list = field(default_factory=list)  # real overlap grids (optional)


def spherical_align(pos1, pos2, Jmax, scale=None, perm=None, invert=True,
                    bw=None, fast=False, keep_grids=False, soft=None,
                    wigner="jacobi"):
    """``BaseSphericalAlignment.align`` (``sphericalAlignment.py:160-188``)
    followed by ``refine`` (``:118-127``).

    ``bw`` overrides the SOFT bandwidth (default Jmax+1 as in ``setJ``,
    ``:44-55``); coefficients are zero-padded to it.  Returns a
    ``SphericalResult`` whose (dist, X1, X2) equals the reference's tuple.
    """
    pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
    pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
    if scale is None:
        scale = (average_separation(pos1) + average_separation(pos2)) / 6
    if perm is None:
        perm = [np.arange(len(pos1))]
    bw = Jmax + 1 if bw is None else bw
    if soft is None:
        soft = SOFTOracle(bw, method=wigner)
    X1 = pos1 - pos1.mean(0)
    X2 = pos2 - pos2.mean(0)

    res = SphericalResult(0.0, X1, X2, None, False, None)
    best = None
    orients = [False, True] if invert else [False]
    for inv in orients:
        X2o = -X2 if inv else X2
        I = so3_coeffs_species(X1, X2o, perm, Jmax, scale, fast)
        Ip = I if bw == Jmax + 1 else pad_coeffs(I, bw)
        grid = soft.iSOFT(Ip)
        ind, find = find_max(grid)
        R0 = soft.indtoEuler(find)
        R, opt = max_overlap(R0, I.conj(), Jmax)
        res.ind.append(tuple(int(i) for i in ind))
        res.fun.append(float(opt.fun))
        if keep_grids:
            res.grids.append(grid)
        # the inverted result replaces only on a strictly smaller objective
        if best is None or opt.fun < best[1]:
            best = (R, opt.fun, inv)
    R, _, inv = best
    X2o = -X2 if inv else X2
    X2R = X2o.dot(euler_matrix(*R))
    _, p = find_best_permutation(X1, X2R, perm)
    dist, MR = kabsch_quaternion(X1, X2R[p])
    res.dist = dist
    res.X2 = X2R[p].dot(MR.T)
    res.perm = p
    res.inverted = inv
    res.R = R
    return res


# ---------------------------------------------------------------------------
# Harmonic-oscillator basis   (SphericalHarmonicAlign,
#                              fastoverlap/sphericalAlignment.py:276-391,
#                              fastoverlap/utils.py:408-427)
# ---------------------------------------------------------------------------


def harmonic_norm(n, l, r0):
    """``norm_harmonicBasis`` (``fastoverlap/utils.py:408-412``)."""
    from scipy.special import gamma
    return (2 * math.factorial(n) * r0 ** (-2 * l - 3) / gamma(1.5 + n + l)) ** 0.5


def harmonic_coeffs(n, l, r0):
    """``coeffs_harmonicBasis`` (``fastoverlap/utils.py:414-427``): the
    coefficients of r^l L_n^{l+1/2}(r^2) in powers of r (scipy's genlaguerre
    coefficients at s = l, l+2, ...), times the normalisation."""
    from scipy.special import genlaguerre
    out = np.zeros(2 * n + l + 1)
    out[l::2] = genlaguerre(n, l + 0.5).coeffs[::-1]
    return out * harmonic_norm(n, l, r0)


def harmonic_coeff_table(nmax, lmax, r0):
    """``SphericalHarmonicAlign.HarmCoeffs`` (``sphericalAlignment.py:320-326``)."""
    c = np.zeros((nmax + 1, lmax + 1, 2 * nmax + lmax + 1))
    for n in range(nmax + 1):
        for l in range(lmax + 1):
            c[n, l, :2 * n + l + 1] = harmonic_coeffs(n, l, r0)
    return c


def radial_integral_harmonic(s, l, rj, sigma, r0):
    """4 pi int_0^inf exp(-(r^2 + rj^2)/2 sigma^2) exp(-r^2/2 r0^2) i_l(r rj/sigma^2)
    r^s r^2 dr in the reference's closed form (``sphericalAlignment.py:287-300``;
    its docstring omits the 4 pi of the Gaussian's multipole expansion)."""
    from scipy.special import gamma, hyp1f1
    return ((2.0 ** (3 + s - l) * pi ** 3) ** 0.5 * rj ** l
            * (sigma ** -2 + r0 ** -2) ** (-0.5 * (3 + s + l)) * sigma ** (-2 * l)
            * hyp1f1(0.5 * (3 + s + l), 1.5 + l,
                     0.5 * rj ** 2 * r0 ** 2 / (r0 ** 2 * sigma ** 2 + sigma ** 4))
            * gamma(0.5 * (3 + l + s)) / gamma(1.5 + l)) * np.exp(-0.5 * rj ** 2 * sigma ** -2)


def harmonic_coeff(pos, Jmax, nmax, scale, r0):
    """c[n, l, m % (2J+1)] of a set of Gaussians of width ``scale``
    (``SphericalHarmonicAlign.calcHarmCoeff``, ``sphericalAlignment.py:340-356``)."""
    pos = np.atleast_2d(np.asarray(pos, float))
    theta, phi, r = calc_theta_phi_r(pos)
    Y = sph_harm_table(Jmax, theta, phi)
    J = np.arange(Jmax + 1)
    dslj = radial_integral_harmonic(np.arange(2 * nmax + Jmax + 1)[:, None, None],
                                    J[None, :, None], r[None, None, :], scale, r0)
    dnlj = np.einsum("nls,slj->nlj", harmonic_coeff_table(nmax, Jmax, r0), dslj)
    return np.einsum("nlj,lmj->nlm", dnlj, Y.conj())


def harmonic_so3_coeffs(X1, X2, perm, Jmax, nmax, scale, r0, basis="python"):
    """``calcSO3Coeffs`` summed over groups (``sphericalAlignment.py:358-361``,
    ``:176``): I[l, m1, m2] = sum_groups sum_n conj(c1[n,l,m1]) c2[n,l,m2];
    ``basis="fortran"``: the coefficients of ``HARMONICCOEFFS`` as summed by
    ``ALIGNHARM`` (``fastoverlap/f90/fastclusters.f90:313-331``)."""
    if basis == "fortran":
        def cf(X):
            return harmonic_coeff_fortran(X, Jmax, nmax, r0, scale)
    else:
        def cf(X):
            return harmonic_coeff(X, Jmax, nmax, scale, r0)
    return sum(np.einsum("nlm,nlo->lmo", cf(X1[p]).conj(), cf(X2[p])) for p in perm)


def harmonic_align(pos1, pos2, Jmax, nmax, scale=None, r0=1.0, perm=None, invert=True,
                   soft=None, basis="python"):
    """``SphericalHarmonicAlign.align`` = ``BaseSphericalAlignment.align``
    (``sphericalAlignment.py:160-188``) with the harmonic-basis coefficients,
    then ``refine`` (``:118-127``); the inverted orientation's coefficients
    are those of -X2 and replace on a strictly smaller objective.
    ``basis="fortran"`` uses ``HARMONICCOEFFS`` (r0 = HWIDTH, scale =
    KWIDTH) with the same search."""
    pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
    pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
    if scale is None:
        scale = (average_separation(pos1) + average_separation(pos2)) / 6
    if perm is None:
        perm = [np.arange(len(pos1))]
    soft = SOFTOracle(Jmax + 1) if soft is None else soft
    X1 = pos1 - pos1.mean(0)
    X2 = pos2 - pos2.mean(0)
    res = SphericalResult(0.0, X1, X2, None, False, None)
    best = None
    for inv in ([False, True] if invert else [False]):
        X2o = -X2 if inv else X2
        I = harmonic_so3_coeffs(X1, X2o, perm, Jmax, nmax, scale, r0, basis)
        ind, find = find_max(soft.iSOFT(I))
        R, opt = max_overlap(soft.indtoEuler(find), I.conj(), Jmax)
        res.ind.append(tuple(int(i) for i in ind))
        res.fun.append(float(opt.fun))
        if best is None or opt.fun < best[1]:
            best = (R, opt.fun, inv)
    R, _, inv = best
    X2o = -X2 if inv else X2
    X2R = X2o.dot(euler_matrix(*R))
    _, p = find_best_permutation(X1, X2R, perm)
    dist, MR = kabsch_quaternion(X1, X2R[p])
    res.dist, res.X2, res.perm, res.inverted, res.R = dist, X2R[p].dot(MR.T), p, inv, R
    return res


def harmonic_nl(nmax, L, rj, sigma, r0):
    """``HARMONICNL`` (``fastoverlap/f90/fastclusters.f90:492-540``): RET[n, l]
    = 4 pi int g_nl(r) exp(-(r^2 + rj^2)/2 sigma^2) i_l(r rj/sigma^2) r^2 dr for
    the normalised oscillator functions
    g_nl = N (r/r0)^l exp(-r^2/2 r0^2) L_n^{l+1/2}(r^2/r0^2), by the
    reference's recurrence over n (degrees up to L - 2n; row 1 to L - 2)."""
    ret = np.zeros((nmax + 1, L + 1))
    rs = 1.0 / (r0 ** 2 + sigma ** 2)
    ret[0, 0] = (math.sqrt(2.0 * math.sqrt(pi) * (r0 * rs) ** 3) * sigma ** 3
                 * math.exp(-0.5 * rj ** 2 * rs) * 4 * pi)
    f0 = math.sqrt(2.0) * r0 * rj * rs
    for j in range(1, L + 1):
        ret[0, j] = f0 / math.sqrt(1.0 + 2.0 * j) * ret[0, j - 1]
    for n in range(1, nmax + 1):
        for j in range(0, (L - 2 if n == 1 else L - 2 * n) + 1):
            v = (math.sqrt(n + j + 0.5) * ret[n - 1, j]
                 - (2.0 * j + 3.0) * sigma ** 2 / rj / r0 * ret[n - 1, j + 1]
                 - math.sqrt(n + j + 1.5) * ret[n - 1, j + 2])
            if n >= 2:
                v += math.sqrt(n - 1.0) * ret[n - 2, j + 2]
            ret[n, j] = v / math.sqrt(n)
    return ret


def harmonic_coeff_fortran(pos, Jmax, nmax, hwidth, kwidth):
    """``HARMONICCOEFFS`` (``fastoverlap/f90/fastclusters.f90:654-687``):
    c[n, l, m % (2J+1)] = sum_j RET_j[n, l] conj(Y_l^m(j))."""
    pos = np.atleast_2d(np.asarray(pos, float))
    theta, phi, r = calc_theta_phi_r(pos)
    Y = sph_harm_table(Jmax, theta, phi)
    D = np.stack([harmonic_nl(nmax, Jmax + 2 * nmax, rj, kwidth, hwidth)[:, :Jmax + 1]
                  for rj in r], axis=-1)
    return np.einsum("nlj,lmj->nlm", D, Y.conj())


def harmonic_similarity(c1s, c2s, Jmax, soft=None):
    """``CALCSIMILARITY`` (``fastoverlap/f90/fastclusters.f90:790-818``) of
    per-group coefficient lists: (sum |I|^2, max of the real grid of
    iSOFT(I) at bandwidth Jmax + 1) with I = sum_groups sum_n conj(c1) c2
    (``DOTHARMONICCOEFFSPERM``, ``:764-788``)."""
    I = sum(np.einsum("nlm,nlo->lmo", a.conj(), b) for a, b in zip(c1s, c2s))
    soft = SOFTOracle(Jmax + 1) if soft is None else soft
    return float((np.abs(I) ** 2).sum()), float(soft.iSOFT(I).real.max())


def harmonic_compare_list(coords, Jmax, nmax, hwidth, kwidth, perm=None):
    """``SphericalHarmonicAlignFortran.compareList`` (``sphericalAlignment.py:
    622-663``, ``CALCOVERLAPMATRICES``): centred structures, symmetric
    matrices of sum |I|^2 and maximum overlap, and their geometric-mean
    normalisations."""
    coords = np.array(coords, float)
    coords -= coords.mean(1)[:, None, :]
    n, N = coords.shape[:2]
    perm = [np.arange(N)] if perm is None else perm
    cs = [[harmonic_coeff_fortran(c[p], Jmax, nmax, hwidth, kwidth) for p in perm]
          for c in coords]
    soft = SOFTOracle(Jmax + 1)
    avg, mx = np.zeros((n, n)), np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            avg[i, j], mx[i, j] = harmonic_similarity(cs[i], cs[j], Jmax, soft)
            avg[j, i], mx[j, i] = avg[i, j], mx[i, j]
    da, dm = avg.diagonal(), mx.diagonal()
    return avg, mx, avg / np.sqrt(np.outer(da, da)), mx / np.sqrt(np.outer(dm, dm))


# ---------------------------------------------------------------------------
# Multi-peak search   (fastoverlap/utils.py:347-396,
#                      fastoverlap/sphericalAlignment.py:196-235)
# ---------------------------------------------------------------------------


def gaussian_nd(x, A, mu, *alphax0):
    """A exp(-sum_{i<=j} s_ij d_i d_j) + mu with d = x - x0
    (``fastoverlap/utils.py:347-354``): ``alphax0`` holds the upper triangle
    of s (row-major) followed by the centre x0; ``x`` is (dim, K)."""
    x = np.atleast_2d(x)
    dim = len(x)
    sig = np.zeros((dim, dim))
    sig[np.triu_indices(dim)] = alphax0[:-dim]
    d = x - np.asarray(alphax0[-dim:], float)[:, None]
    return A * np.exp(-np.einsum("ik,jk,ij->k", d, d, sig)) + mu


def fit_peak(f, ind, n=2):
    """Least-squares Gaussian on the (2n+1)^dim block around ``ind``, indices
    wrapped periodically (``fastoverlap/utils.py:356-365``); scipy curve_fit
    from A = peak value, mu = 0, s = identity, x0 = 0."""
    from scipy.optimize import curve_fit
    dim = f.ndim
    block = f[np.ix_(*[np.arange(i - n, i + n + 1) % sz for i, sz in zip(ind, f.shape)])]
    pts = np.indices((2 * n + 1,) * dim).reshape((dim, -1)) - n
    p0 = ([block[(n,) * dim], 0.0]
          + [1.0 if i == j else 0.0 for i in range(dim) for j in range(i, dim)]
          + [0.0] * dim)
    popt, pcov = curve_fit(gaussian_nd, pts, block.ravel(), p0=p0)
    return popt, pcov


def find_peaks(a, npeaks=10, width=2):
    """Subtractive Gaussian peak search (``fastoverlap/utils.py:367-396``):
    from f = a - min(a), repeat npeaks times: argmax (first occurrence), fit
    a Gaussian around it, record its centre, subtract the fitted Gaussian
    (background mu included, no wrap) from the whole array; stop at the
    first failed fit.  Without any peak, the interpolated maximum.
    Returns (peaks (k, dim) fractional indices, amplitudes, backgrounds,
    widths, residual array)."""
    f = np.array(a, dtype=float)
    f -= f.min()
    dim = f.ndim
    idx = np.indices(f.shape).reshape((dim, -1))
    peaks, amp, mean, sigma = [], [], [], []
    for _ in range(npeaks):
        ind = np.unravel_index(f.argmax(), f.shape)
        try:
            popt = fit_peak(f, ind, width)[0]
        except (RuntimeError, ValueError):
            break
        peaks.append(popt[-dim:] + np.asarray(ind))
        amp.append(popt[0])
        mean.append(popt[1])
        with np.errstate(divide="ignore", invalid="ignore"):
            sigma.append((2 * popt[2:-dim]) ** -0.5)
        popt[-dim:] += np.asarray(ind)
        f[tuple(idx)] -= gaussian_nd(idx, *popt)
    peaks = np.array(peaks)
    if len(peaks) == 0:
        peaks = find_max(a)[1][None, :]
        amp.append(np.max(a))
        mean.append(0.0)
        sigma.append(np.nan)
    return peaks, amp, mean, sigma, f


def find_rotations(soft, Ilmm, nrot=10, width=2):
    """``fastoverlap/sphericalAlignment.py:196-205``: Euler angles of the
    peaks of the real overlap grid (window widened until a fit succeeds)."""
    overlap = soft.iSOFT(Ilmm).real
    peaks = []
    while len(peaks) == 0:
        peaks, amp, mean, sigma, f = find_peaks(overlap, nrot, width)
        width += 1
    return np.atleast_2d(soft.indtoEuler(peaks)), peaks


def refine_rotation(X1, X2, R, perm):
    """``fastoverlap/sphericalAlignment.py:118-127``: rotate X2 by
    EulerM(R), per-group LAP, quaternion fit.  Returns (dist, X2 aligned,
    permutation)."""
    X2R = X2.dot(euler_matrix(*R))
    _, p = find_best_permutation(X1, X2R, perm)
    dist, MR = kabsch_quaternion(X1, X2R[p])
    return dist, X2R[p].dot(MR.T), p


def spherical_malign(pos1, pos2, Jmax, scale, perm=None, invert=True, bw=None, nrot=10,
                     soft=None, fast=True):
    """``BaseSphericalAlignment.malign`` (``sphericalAlignment.py:206-235``):
    refine from every peak of each orientation's grid, keep the smallest
    distance.  The inverted branch refines -X2 (the intended semantics; the
    reference passes X2 there).  Returns (dist, X1, X2, perm, inverted,
    peaks per orientation)."""
    pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
    pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
    if perm is None:
        perm = [np.arange(len(pos1))]
    bw = Jmax + 1 if bw is None else bw
    soft = SOFTOracle(bw) if soft is None else soft
    X1 = pos1 - pos1.mean(0)
    X2 = pos2 - pos2.mean(0)
    best, allpeaks = None, []
    for inv in ([False, True] if invert else [False]):
        X2o = -X2 if inv else X2
        I = so3_coeffs_species(X1, X2o, perm, Jmax, scale, fast)
        Ip = I if bw == Jmax + 1 else pad_coeffs(I, bw)
        Rs, peaks = find_rotations(soft, Ip, nrot)
        allpeaks.append(peaks)
        for R in Rs:
            d, X2a, p = refine_rotation(X1, X2o, R, perm)
            # the inverted result replaces only on a strictly smaller distance
            if best is None or d < best[0]:
                best = (d, X2a, p, inv)
    d, X2a, p, inv = best
    return d, X1, X2a, p, inv, allpeaks


# ---------------------------------------------------------------------------
# Periodic alignment                (fastoverlap/periodicAlignment.py)
# ---------------------------------------------------------------------------


def next_fast_len(target: int) -> int:
    """Smallest 2,3,5-smooth integer >= target (``fastoverlap/utils.py:278-313``)."""
    if target <= 6:
        return target
    if not (target & (target - 1)):
        return target
    best = float("inf")
    p5 = 1
    while p5 < target:
        p35 = p5
        while p35 < target:
            q = -(-target // p35)
            p2 = 2 ** ((q - 1).bit_length())
            N = p2 * p35
            if N == target:
                return N
            best = min(best, N)
            p35 *= 3
            if p35 == target:
                return p35
        best = min(best, p35)
        p5 *= 5
        if p5 == target:
            return p5
    return int(min(best, p5))


class PeriodicOracle:
    """``PeriodicAlign`` (``fastoverlap/periodicAlignment.py:341-479``)."""

    def __init__(self, Natoms, boxvec, permlist=None, scale=None, n=None,
                 fshape=None):
        self.Natoms = Natoms
        self.boxvec = np.array(boxvec, float)
        self.dim = 3
        self.perm = ([np.arange(Natoms)] if permlist is None
                     else [np.asarray(p) for p in permlist])
        if scale is None:
            scale = (np.prod(self.boxvec) / Natoms) ** (1.0 / 3) / 3.0
        self.scale = scale
        self.n = int(np.ceil(1.3 * Natoms ** (1.0 / 3.0))) if n is None else n
        ps = np.indices((self.n * 2 + 1,) * 3) - self.n
        self.ks = 2.0 * pi / self.boxvec[:, None, None, None] * ps
        self.absks = np.linalg.norm(self.ks, axis=0)
        shape = np.array(self.absks.shape) * 2 + 1
        self.fshape = (tuple(next_fast_len(int(d)) for d in shape)
                       if fshape is None else tuple(fshape))

    def calc_fourier_coeff(self, pos):
        """``periodicAlignment.py:400-406``: C[p][k] = sum_j exp(-i k.r_j)."""
        out = np.empty((len(self.perm),) + self.absks.shape, complex)
        for p, C in zip(self.perm, out):
            np.exp(-1j * np.tensordot(pos[p, :], self.ks, [1, 0])).sum(0, out=C)
        return out

    def overlap_grid(self, pos1, pos2, Cs=None):
        """``periodicAlignment.py:408-440``: |fftn(sum_p C1 conj(C2) e^{-k^2 s^2})|."""
        if Cs is None:
            C1 = self.calc_fourier_coeff(pos1)
            C2 = self.calc_fourier_coeff(pos2)
        else:
            C1, C2 = Cs
        C = (C1 * C2.conj() * np.exp(-self.absks ** 2 * self.scale ** 2)).sum(0)
        f = np.fft.fftn(C, self.fshape)
        return C, np.abs(f)

    def periodic(self, x):
        return x - np.round(x / self.boxvec) * self.boxvec

    def cost_matrix(self, X1, X2):
        """Periodic Euclidean cost, rows = X1 (intended semantics; the
        reference's ``periodicAlignment.py:94-102`` puts X2 on the rows)."""
        d = X1[:, None, :] - X2[None, :, :]
        d -= np.round(d / self.boxvec) * self.boxvec
        return np.linalg.norm(d, axis=2)

    def hungarian(self, X1, X2):
        """``periodicAlignment.py:104-107`` with intended semantics."""
        _, perm = find_best_permutation(X1, X2, self.perm, self.cost_matrix)
        return np.linalg.norm(self.periodic(X1 - X2[perm])), perm

    def refine(self, x, y, disps, niter=10):
        """``periodicAlignment.py:27-80``."""
        disps = np.atleast_2d(disps)
        cand = [self.hungarian(x, y - d[None, :]) + (d.copy(),) for d in disps]
        dist, saveperm, disp = min(cand, key=lambda t: t[0])
        perm = saveperm
        for _ in range(niter):
            dxs = self.periodic(x - (y - disp)[saveperm])
            disp = disp - dxs.mean(0)
            perm = self.hungarian(x, y - disp[None, :])[1]
            if np.array_equal(saveperm, perm):
                break
            saveperm = perm
        dxs = self.periodic(x - (y - disp)[perm])
        disp = disp - dxs.mean(0)
        pos1 = self.periodic(x.copy())
        pos2 = self.periodic(y[perm] - disp)
        dist = np.linalg.norm(self.periodic(pos1 - pos2))
        return dist, pos1, pos2, perm, disp

    def find_disps(self, fabs, npeaks=1, width=2):
        """``periodicAlignment.py:442-456``: the interpolated maximum
        (npeaks = 1) or the Gaussian-fitted peaks of |f|, times box/F."""
        shape = np.array(fabs.shape)
        if npeaks > 1:
            disps = find_peaks(fabs, npeaks, width)[0]
            return np.atleast_2d(disps) * self.boxvec / shape
        return (find_max(fabs)[1] * self.boxvec / shape)[None, :]

    def align(self, pos1, pos2, Cs=None, npeaks=1, width=2):
        """``periodicAlignment.py:442-460``."""
        pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
        pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
        C, fabs = self.overlap_grid(pos1, pos2, Cs)
        ind, _ = find_max(fabs)
        disps = self.find_disps(fabs, npeaks, width)
        out = self.refine(pos1, pos2, disps)
        return out, tuple(int(i) for i in ind), fabs


# The 48 operations of the cubic point group O_h, in the order of OHOPS
# (fastoverlap/f90/alignutils.f90:830-987): for output axis x, y, z the sign
# and the input axis, so "+y-x+z" maps (x, y, z) to (y, -x, z).
OH_OPS = ("+x+y+z -x-y+z +y+z+x +y-x+z -x-y-z -y+z-x -y+x+z +x+y-z "
          "-y-z+x +y-z-x +z+x+y +z-x-y +z-y+x +z+y-x -y-z-x +y-z+x "
          "-z+x-y -z-x+y -z+y+x -z-y-x +y+z-x -y+z+x +x-y-z -x+y-z "
          "+x-z+y -x-z-y +x+z-y -x+z+y -z-x-y -z+x+y +z-x+y +z+x-y "
          "-y-x-z +y+x-z -x+y+z +x-y+z -y+x-z +y-x-z -z+y-x -z-y+x "
          "+z-y-x +z+y+x -x+z-y +x+z+y -x-z+y +x-z-y +y+x+z -y-x+z").split()


def oh_matrices():
    """Row-major matrices of ``OH_OPS`` (y = R x per atom, OHOPS at
    ``fastoverlap/f90/alignutils.f90:983-987``)."""
    out = np.zeros((48, 3, 3))
    for o, code in enumerate(OH_OPS):
        for r in range(3):
            out[o, r, "xyz".index(code[2 * r + 1])] = -1.0 if code[2 * r] == "-" else 1.0
    return out


def bulk_align(P, pos1, pos2, ndisps=1, ohcell=0, width=2):
    """``BULKFASTOVERLAP::ALIGN1`` + ``ALIGNCOEFFS``
    (``fastoverlap/f90/fastbulk.f90:414-565``) over the Python building
    blocks of ``PeriodicOracle`` ``P``: candidate displacements of the
    overlap of (pos1, pos2) (``find_disps``), each refined on its own
    (``refine`` with one displacement), the first strictly smallest distance
    kept.  ``ohcell=1`` repeats the refinement for the 48 images
    ``pos2 @ R.T`` from the displacements of the untransformed pair, as the
    loop at ``fastbulk.f90:458-481`` does (it transforms FCOEFFA into
    FCOEFFDUMMYA but hands ALIGNCOEFFS the untransformed FCOEFFA);
    ``ohcell=2`` takes each image's displacements from its own overlap.
    Returns ((dist, X1, X2, perm, disp), operation)."""
    pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
    pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
    mats = oh_matrices() if ohcell else np.eye(3)[None]
    if ohcell != 2:
        disps0 = P.find_disps(P.overlap_grid(pos1, pos2)[1], ndisps, width)
    best = None
    for o, R in enumerate(mats):
        Y = pos2 @ R.T
        disps = disps0 if ohcell != 2 else P.find_disps(P.overlap_grid(pos1, Y)[1], ndisps,
                                                         width)
        for d in disps:
            out = P.refine(pos1, Y, d[None, :])
            if best is None or out[0] < best[0][0]:
                best = (out, o)
    return best


def bulk_defaults(natoms, boxvec):
    """``CALCDEFAULTS`` (``fastoverlap/f90/fastbulk.f90:253-277``):
    (kernel width, NWAVE, NFSPACE)."""
    boxvec = np.asarray(boxvec, float)
    kw = (np.prod(boxvec) / natoms) ** (1.0 / 3.0) / 3.0
    nwave = int(math.ceil(2 * 3.14159265359 / boxvec.min() * (1.5 / kw)))
    nf = (next_fast_len(4 * nwave + 3) if 2 * nwave + 1 <= 200
          else 2 ** int(math.ceil(math.log2(4 * nwave + 3))))
    return kw, nwave, nf


def bulk_align_group(coords1, coords2=None, boxvec=None, perm=None, ndisps=1, sym=True):
    """``ALIGNGROUP`` (``fastoverlap/f90/fastbulk.f90:336-412``) at the
    CALCDEFAULTS width and k-cube: structure j of coords2 aligned onto
    structure i of coords1 for i >= j (sym, overlap of j formed from
    coords1[j]) or every (i, j).  Returns (distmat [n1, n2], aligned
    [n1, n2, N, 3]); cells sym leaves out are zero."""
    coords1 = np.asanyarray(coords1, float)
    n1, natoms = coords1.shape[:2]
    coords2 = coords1 if (sym or coords2 is None) else np.asanyarray(coords2, float)
    n2 = coords2.shape[0]
    kw, nwave, nf = bulk_defaults(natoms, boxvec)
    P = PeriodicOracle(natoms, boxvec, perm, scale=kw, n=nwave, fshape=(nf,) * 3)
    dist = np.zeros((n1, n2))
    aligned = np.zeros((n1, n2, natoms, 3))
    for j in range(n2):
        for i in range(j if sym else 0, n1):
            src = coords1[j] if sym else coords2[j]
            fabs = P.overlap_grid(coords1[i], src)[1]
            best = None
            for d in P.find_disps(fabs, ndisps, 2):
                out = P.refine(coords1[i], coords2[j], d[None, :])
                if best is None or out[0] < best[0]:
                    best = out
            dist[i, j] = best[0]
            aligned[i, j] = best[2]
    return dist, aligned


PINNED = {
    "makeweights": "quadrature exactness on Wigner-d orthogonality (test_oracle)",
    "wigner_d_jacobi": "small-l closed forms, orthogonality, unitarity",
    "wigner_d_recurrence": "Jacobi closed form to l = 40 on the grid-128 betas; "
                           "quadrature orthonormality to l = 63",
    "SOFTOracle.iSOFT": "round trip SOFT(iSOFT(f)) = f; delta coefficients",
    "SOFTOracle.SOFT": "round trip with iSOFT",
    "pad_coeffs": "round trip at bw > L+1; direct-sum grid values; delta at bw = 2(L+1)",
    "sph_harm_table": "scipy routine; quadrature orthonormality; Y00",
    "radial_kernel": "scipy spherical_in; limits at x = 0",
    "so3_coeffs": "brute-force overlap identity; Hermitian; inversion (-1)^l",
    "find_max": "numpy argmax semantics; parabola exact on a quadratic",
    "energy_gradient": "finite differences; value equals the grid identity",
    "max_overlap": "exact copy -> true rotation; stationary local minimum; Nelder-Mead agrees",
    "find_best_permutation": "brute force over all permutations, N <= 7",
    "kabsch_quaternion": "SVD Kabsch; exact rotation gives 0",
    "average_separation": "cubic lattice closed form; k-d tree nearest neighbours",
    "spherical_align": "LJ38 known answer 1.4767 (Jmax=14); exact rotation -> 0",
    "PeriodicOracle.overlap_grid": "peak at the known translation; direct DFT",
    "PeriodicOracle.hungarian": "brute force (two groups); exact shifted copy; BLJ256 1.559",
    "PeriodicOracle.refine": "exact translation+permutation -> 0; BLJ256 1.559",
    "gaussian_nd": "closed form at known points (upper-triangle quadratic form)",
    "find_peaks": "sums of known Gaussians: every centre, amplitude and the "
                  "argmax sequence recovered; brute-force argmax on small grids",
    "spherical_malign": "exact rotation -> 0; agrees with align on easy pairs",
    "PeriodicOracle.find_disps": "known translation recovered with npeaks > 1",
    "oh_matrices": "the reference's OHOPS table and order (fixture from its DATA "
                   "statement); 48 distinct signed permutations closed under products",
    "bulk_align": "reduces to align (one displacement, no Oh search); translated copy -> 0 "
                  "at operation 0; octahedral image -> 0 at the inverse operation (ohcell=2)",
    "bulk_defaults": "closed form for BLJ256",
    "bulk_align_group": "every cell equals bulk_align of that pair; sym leaves i < j zero",
    "harmonic_norm": "basis orthonormality by quadrature at r0 = 1",
    "harmonic_coeffs": "basis orthonormality by quadrature; Laguerre closed form",
    "harmonic_coeff_table": "rows equal harmonic_coeffs",
    "radial_integral_harmonic": "numerical quadrature of the defining integral",
    "harmonic_coeff": "single atom on the z axis: only m = 0, equal to the radial integrals "
                      "times Y_l0; rotation covariance of |c| per (n, l)",
    "harmonic_so3_coeffs": "Hermitian identity of a pair with itself; inversion (-1)^l",
    "harmonic_align": "exact rotated, permuted copy -> 0; inverted copy -> 0 and inverted",
    "harmonic_nl": "quadrature of the defining integral with the normalised oscillator "
                   "functions (Laguerre polynomials from scipy)",
    "harmonic_coeff_fortran": "single atom on the z axis: only m = 0, RET times Y_l0; "
                              "rotation invariance of the power per (n, l)",
    "harmonic_similarity": "sum |I|^2 invariant under rotating either structure (unitary "
                           "D^l); maximum equals the brute-force grid maximum",
    "harmonic_compare_list": "symmetric, unit normalised diagonal, rotated copy has the "
                             "self-overlap norm",
}
