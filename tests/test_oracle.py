"""Pinning of the FP64 oracle (CPU).  Each check is chosen so that a plausible
mistake -- a dropped term, a wrong sign or index, a transposed operand --
fails it: known answers from the reference's own example data, closed forms,
library routines, invariants and brute force on tiny inputs."""

import itertools
import json
import math
import os

import numpy as np
import pytest
from scipy.fft import next_fast_len as scipy_next_fast_len
from scipy.special import spherical_in

from oracle import fo_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- Wigner-d -----------------------------------------------------------------

def test_wigner_small_l_closed_forms():
    b = np.array([0.3, 1.1, 2.7])
    d = O.wigner_d_jacobi(1, b, normalised=False)
    n = 3
    c, s = np.cos(b), np.sin(b)
    # textbook d^1 (z-y-z convention): d^1_{11} = (1+c)/2, d^1_{10} = -s/sqrt2,
    # d^1_{1,-1} = (1-c)/2, d^1_{00} = c, d^1_{01} = s/sqrt2
    np.testing.assert_allclose(d[1, 1, 1], (1 + c) / 2, atol=1e-15)
    np.testing.assert_allclose(d[1, 1, 0], -s / math.sqrt(2), atol=1e-15)
    np.testing.assert_allclose(d[1, 1, -1 % n], (1 - c) / 2, atol=1e-15)
    np.testing.assert_allclose(d[1, 0, 0], c, atol=1e-15)
    np.testing.assert_allclose(d[1, 0, 1], s / math.sqrt(2), atol=1e-15)
    np.testing.assert_allclose(d[0, 0, 0], 1.0, atol=1e-15)


def test_wigner_unitary():
    L = 12
    d = O.wigner_d_jacobi(L, np.array([0.83]), normalised=False)[..., 0]
    n = 2 * L + 1
    for l in (1, 5, 12):
        idx = [m % n for m in range(-l, l + 1)]
        M = d[l][np.ix_(idx, idx)]
        np.testing.assert_allclose(M @ M.T, np.eye(2 * l + 1), atol=1e-12)


def test_wigner_quadrature_orthogonality():
    # SOFT quadrature: sum_k w_k dt^l dt^l' (b_k) = delta_{ll'} for |m| <= l
    bw = 10
    soft = O.SOFTOracle(bw)
    Ds, w = soft.Ds, soft.weights
    n = 2 * bw - 1
    for m1, m2 in [(0, 0), (2, -1), (3, 3), (-4, 1)]:
        l0 = max(abs(m1), abs(m2))
        G = np.array([[np.sum(w * Ds[l, m1 % n, m2 % n] * Ds[lp, m1 % n, m2 % n])
                       for lp in range(l0, bw)] for l in range(l0, bw)])
        np.testing.assert_allclose(G, np.eye(bw - l0), atol=1e-12)


def test_wigner_recurrence_matches_jacobi():
    bw = 31
    b = O.soft_grid(bw)[1]
    np.testing.assert_allclose(O.wigner_d_recurrence(30, b), O.wigner_d_jacobi(30, b),
                               atol=5e-12)


# ---- SOFT ---------------------------------------------------------------------------

def _random_flmm(rng, bw):
    f = np.zeros((bw, 2 * bw - 1, 2 * bw - 1), complex)
    n = 2 * bw - 1
    for l in range(bw):
        for m1 in range(-l, l + 1):
            for m2 in range(-l, l + 1):
                f[l, m1 % n, m2 % n] = rng.normal() + 1j * rng.normal()
    return f


def test_soft_round_trip():
    rng = np.random.default_rng(0)
    soft = O.SOFTOracle(6)
    f = _random_flmm(rng, 6)
    np.testing.assert_allclose(soft.SOFT(soft.iSOFT(f)), f, atol=1e-11)


def test_isoft_delta_coefficient():
    bw = 5
    soft = O.SOFTOracle(bw)
    f = np.zeros((bw, 2 * bw - 1, 2 * bw - 1), complex)
    l, m1, m2 = 3, 2, -1
    f[l, m1, m2] = 1.0
    g = soft.iSOFT(f)
    a, b, y = soft.a, soft.b, soft.y
    expect = (np.exp(1j * m1 * a)[:, None, None] * soft.Ds[l, m1, m2][None, :, None]
              * np.exp(1j * m2 * y)[None, None, :])
    np.testing.assert_allclose(g, expect, atol=1e-13)


# ---- coefficients -----------------------------------------------------------------

def test_sph_harm_basics():
    th = np.array([0.4, 1.3, 2.9])
    ph = np.array([0.1, -2.0, 3.0])
    Y = O.sph_harm_table(3, th, ph)
    np.testing.assert_allclose(Y[0, 0], 1 / math.sqrt(4 * math.pi))
    # Y_1^0 = sqrt(3/4pi) cos(theta); Y_1^{-1} = (-1)^1 conj(Y_1^1)
    np.testing.assert_allclose(Y[1, 0], math.sqrt(3 / (4 * math.pi)) * np.cos(th))
    np.testing.assert_allclose(Y[1, -1 % 7], -np.conj(Y[1, 1]))
    # Y_1^1 = -sqrt(3/8pi) sin(theta) e^{i phi} (Condon-Shortley)
    np.testing.assert_allclose(Y[1, 1], -math.sqrt(3 / (8 * math.pi)) * np.sin(th)
                               * np.exp(1j * ph))


def test_sph_harm_orthonormal_by_quadrature():
    L = 6
    xg, wg = np.polynomial.legendre.leggauss(2 * L + 2)
    th = np.arccos(xg)
    nphi = 4 * L + 4
    ph = 2 * np.pi * np.arange(nphi) / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    W = np.repeat(wg, nphi) * (2 * np.pi / nphi)
    Y = O.sph_harm_table(L, T.ravel(), P.ravel()).reshape((L + 1) * (2 * L + 1), -1)
    G = (Y * W) @ Y.conj().T
    keep = np.abs(np.diag(G)) > 0.5
    np.testing.assert_allclose(G[np.ix_(keep, keep)], np.eye(keep.sum()), atol=1e-12)
    assert keep.sum() == (L + 1) ** 2


def test_radial_kernel_matches_spherical_in():
    r1 = np.array([0.0, 0.3, 1.2, 2.0])
    r2 = np.array([0.5, 1.0, 1.7, 2.2])
    s = 0.4
    il = O.radial_kernel(10, r1, r2, s)
    x = r1[:, None] * r2[None] / s ** 2 / 2
    for l in range(11):
        expect = spherical_in(l, x) * np.exp(-(r1[:, None] ** 2 + r2[None] ** 2) / 4 / s ** 2)
        np.testing.assert_allclose(il[l], expect, rtol=1e-12, atol=1e-300)
    # x = 0 limit: i_0(0) = 1, i_l(0) = 0
    assert il[0, 0, 0] == pytest.approx(np.exp(-r2[0] ** 2 / 4 / s ** 2))
    assert np.all(il[1:, 0, :] == 0)


def _cloud(rng, n, R=1.4):
    X = rng.normal(size=(n, 3))
    X *= R / np.linalg.norm(X, axis=1).max()
    return X - X.mean(0)


def test_coeffs_hermitian_and_inversion():
    rng = np.random.default_rng(3)
    X1, X2 = _cloud(rng, 9), _cloud(rng, 9)
    L = 6
    n = 2 * L + 1
    I = O.so3_coeffs(X1, X2, L, 0.35)
    for l in range(L + 1):
        for m1 in range(-l, l + 1):
            for m2 in range(-l, l + 1):
                assert I[l, -m1 % n, -m2 % n] == pytest.approx(
                    (-1) ** (m1 + m2) * np.conj(I[l, m1 % n, m2 % n]), abs=1e-13)
    Iinv = O.so3_coeffs(X1, -X2, L, 0.35)
    np.testing.assert_allclose(Iinv, ((-1.0) ** np.arange(L + 1))[:, None, None] * I,
                               atol=1e-13)
    np.testing.assert_allclose(O.so3_coeffs(X1, X2, L, 0.35, fast=True), I, atol=1e-13)


def test_coeffs_brute_force_overlap_identity():
    # Re sum conj(I) D(R) equals the direct Gaussian overlap of X1 with X2
    # rotated as refine() rotates it (X2 . EulerM(R)), up to truncation in l.
    rng = np.random.default_rng(7)
    X1, X2 = _cloud(rng, 6, 1.0), _cloud(rng, 6, 1.0)
    s = 0.5
    L = 22
    I = O.so3_coeffs(X1, X2, L, s)
    for R in [np.array([0.3, 1.2, -0.7]), np.array([2.0, 0.4, 1.0])]:
        f, _ = O.energy_gradient(R, I.conj(), L)
        brute = O.brute_overlap(X1, X2.dot(O.euler_matrix(*R)), s)
        assert -f == pytest.approx(brute, rel=1e-9)


# ---- peak search / refinement ------------------------------------------------------

def test_find_max_parabola_exact():
    n = 12
    i = np.arange(n)
    c = np.array([4.3, 7.8, 2.2])
    g = -sum(((np.meshgrid(i, i, i, indexing="ij")[k] - c[k]) ** 2) for k in range(3)) + 100
    ind, f = O.find_max(g)
    np.testing.assert_array_equal(ind, [4, 8, 2])
    np.testing.assert_allclose(f, c, atol=1e-12)


def test_find_max_complex_lexicographic_first_occurrence():
    a = np.zeros((2, 2, 2), complex)
    a[0, 1, 1] = 3 + 0j
    a[1, 0, 0] = 3 + 1e-9j
    a[1, 1, 0] = 3 + 1e-9j
    ind, _ = O.find_max(a)
    np.testing.assert_array_equal(ind, [1, 0, 0])


def test_energy_gradient_finite_difference_and_recurrence():
    rng = np.random.default_rng(1)
    L = 8
    X = rng.normal(size=(L + 1, 2 * L + 1, 2 * L + 1)) + 1j * rng.normal(
        size=(L + 1, 2 * L + 1, 2 * L + 1))
    rot = np.array([0.7, 1.1, 2.3])
    f, g = O.energy_gradient(rot, X, L)
    f2, g2 = O.energy_gradient_recurrence(rot, X, L)
    h = 1e-6
    fd = [(O.energy_gradient(rot + h * e, X, L)[0] - O.energy_gradient(rot - h * e, X, L)[0])
          / (2 * h) for e in np.eye(3)]
    assert f == pytest.approx(f2, rel=1e-12)
    np.testing.assert_allclose(g, g2, rtol=1e-11)
    np.testing.assert_allclose(g, fd, rtol=1e-7)


def test_lap_brute_force():
    rng = np.random.default_rng(5)
    for n in (3, 5, 7):
        X1, X2 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        _, perm = O.find_best_permutation(X1, X2, [np.arange(n)])
        cost = ((X1[:, None] - X2[None]) ** 2).sum(-1)
        best = min(itertools.permutations(range(n)),
                   key=lambda p: cost[np.arange(n), list(p)].sum())
        np.testing.assert_array_equal(perm, best)


def test_kabsch_matches_svd_and_exact_rotation():
    rng = np.random.default_rng(9)
    X = rng.normal(size=(20, 3))
    X -= X.mean(0)
    Rm = O.euler_matrix(0.4, 1.9, -1.2)
    dist, MR = O.kabsch_quaternion(X, X.dot(Rm.T))
    assert dist < 1e-6
    Y = rng.normal(size=(20, 3))
    Y -= Y.mean(0)
    dist, MR = O.kabsch_quaternion(X, Y)
    U, S, Vt = np.linalg.svd(Y.T @ X)
    dsign = np.sign(np.linalg.det(U @ Vt))
    Rsvd = U @ np.diag([1, 1, dsign]) @ Vt
    assert dist == pytest.approx(np.linalg.norm(X - Y @ Rsvd), rel=1e-10)
    np.testing.assert_allclose(np.linalg.norm(X - Y.dot(MR.T)), dist, rtol=1e-10)


# ---- end to end: the reference's known answers ------------------------------------

def test_lj38_known_answer(lj38):
    # fastoverlap/sphericalAlignment.py:680-693: "distance should = 1.4767"
    p1, p2 = lj38
    r = O.spherical_align(p1, p2, 14, 0.3, fast=True)
    # the reference prints the answer to 4 decimals; its L-BFGS-B end point
    # only enters through the permutation, so agreement to 1e-4 pins it
    assert r.dist == pytest.approx(1.4767, abs=1e-4)
    assert np.linalg.norm(r.X1 - r.X2) == pytest.approx(r.dist, rel=1e-10)


def test_random_cloud_exact_rotation_gives_zero():
    # fastoverlap/sphericalAlignment.py:711-723: "distance should ~ 0"
    rng = np.random.default_rng(11)
    X = rng.normal(size=(30, 3)) * 1.2
    X -= X.mean(0)
    Rm = O.euler_matrix(1.0, 0.7, 2.5)
    Y = X.dot(Rm.T)[rng.permutation(30)]
    r = O.spherical_align(X, Y, 12, 0.3, fast=True)
    assert r.dist < 1e-6


def test_blj256_known_answer(blj256):
    # fastoverlap/periodicAlignment.py:609-633: "distance should = 1.559"
    p1, p2 = blj256
    P = O.PeriodicOracle(256, np.ones(3) * 5.975206329, [np.arange(204), np.arange(204, 256)])
    (dist, X1, X2, perm, disp), ind, _ = P.align(p1, p2)
    assert dist == pytest.approx(1.559, abs=1e-3)
    assert P.fshape == (40, 40, 40) and P.n == 9


def test_periodic_grid_peak_and_corner_padding():
    rng = np.random.default_rng(2)
    box = np.full(3, 5.0)
    X = rng.random((20, 3)) * box
    shift = np.array([1.0, 2.5, 3.75])       # multiples of box/F with F = 20
    P = O.PeriodicOracle(20, box, n=4, fshape=(20, 20, 20))
    C, fabs = P.overlap_grid(X, X + shift)
    ind, find = O.find_max(fabs)
    np.testing.assert_array_equal(ind, np.round(shift / box * 20).astype(int))
    # |fftn| of the corner-padded product equals a direct DFT
    x = np.array([3, 7, 11])
    k = np.arange(9)
    ph = np.exp(-2j * np.pi * np.add.outer(np.add.outer(k * x[0], k * x[1]), k * x[2]) / 20)
    assert abs((C * ph).sum()) == pytest.approx(fabs[tuple(x)], rel=1e-12)


def test_periodic_exact_copy_aligns_to_zero():
    rng = np.random.default_rng(4)
    box = np.full(3, 6.0)
    X = rng.random((40, 3)) * box
    perm = rng.permutation(40)
    Y = (X + np.array([1.1, -2.2, 0.7]))[perm]
    P = O.PeriodicOracle(40, box)
    (dist, X1, X2, p, disp), _, _ = P.align(X, Y)
    assert dist < 1e-9
    np.testing.assert_array_equal(perm[p], np.arange(40))


def test_next_fast_len():
    def smooth(v):
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        return v == 1
    for t in range(7, 400):
        assert O.next_fast_len(t) == next(v for v in itertools.count(t) if smooth(v))
    for t in range(1, 7):
        assert O.next_fast_len(t) == t
    # scipy's rule (which also admits 7 and 11) agrees on 2,3,5-smooth inputs
    assert all(scipy_next_fast_len(O.next_fast_len(t)) == O.next_fast_len(t)
               for t in range(7, 400))


# ---- pinning of the remaining oracle functions ----------------------------------------

def test_wigner_recurrence_matches_jacobi_at_l40_on_c3_grid():
    # the large-degree workload (Jmax 40, grid 128): beta_k = pi (2k+1)/256
    b = O.soft_grid(64)[1]
    np.testing.assert_allclose(O.wigner_d_recurrence(40, b), O.wigner_d_jacobi(40, b),
                               atol=5e-12)


def test_wigner_recurrence_quadrature_orthonormal_to_l63():
    # SOFT quadrature at bw = 64 is exact for products of degree <= 63, so the
    # recurrence's normalised d must be orthonormal in l for every (m1, m2),
    # independent of the Jacobi closed form (which loses accuracy up there)
    bw = 64
    b = O.soft_grid(bw)[1]
    w = O.makeweights(bw)
    D = O.wigner_d_recurrence(63, b)
    n = 127
    for m1, m2 in [(0, 0), (63, -63), (40, 17), (-5, 62), (31, 31), (-62, -1)]:
        l0 = max(abs(m1), abs(m2))
        rows = D[l0:, m1 % n, m2 % n]
        G = (rows * w[None]) @ rows.T
        np.testing.assert_allclose(G, np.eye(64 - l0), atol=1e-11)


def test_pad_coeffs_soft_round_trip_and_grid_identity():
    # pad(I) on a larger grid: SOFT(iSOFT(pad(f))) returns pad(f), and each grid
    # value is the direct sum  sum_l,m1,m2 I^l_{m1m2} dn^l_{m1m2}(b) e^{i(m1 a + m2 y)}
    rng = np.random.default_rng(21)
    L, bw = 4, 8
    f = _random_flmm(rng, L + 1)
    soft = O.SOFTOracle(bw)
    fp = O.pad_coeffs(f, bw)
    assert fp.shape == (bw, 2 * bw - 1, 2 * bw - 1)
    assert np.count_nonzero(fp) == np.count_nonzero(f)
    np.testing.assert_allclose(soft.SOFT(soft.iSOFT(fp)), fp, atol=1e-11)
    g = soft.iSOFT(fp)
    n = 2 * L + 1
    for ia, ib, ig in [(0, 0, 0), (3, 7, 11), (15, 2, 9)]:
        a, b, y = soft.a[ia], soft.b[ib], soft.y[ig]
        d = O.wigner_d_jacobi(L, np.array([b]))[..., 0]
        s = 0j
        for l in range(L + 1):
            for m1 in range(-l, l + 1):
                for m2 in range(-l, l + 1):
                    s += (f[l, m1 % n, m2 % n] * d[l, m1 % n, m2 % n]
                          * np.exp(1j * (m1 * a + m2 * y)))
        assert g[ia, ib, ig] == pytest.approx(s, abs=1e-11)


def test_pad_coeffs_delta_at_twice_the_bandwidth():
    L = 5
    bw = 2 * (L + 1)
    f = np.zeros((L + 1, 2 * L + 1, 2 * L + 1), complex)
    l, m1, m2 = 4, -3, 2
    f[l, m1, m2] = 1.0
    soft = O.SOFTOracle(bw)
    g = soft.iSOFT(O.pad_coeffs(f, bw))
    n = 2 * bw - 1
    expect = (np.exp(1j * m1 * soft.a)[:, None, None] * soft.Ds[l, m1 % n, m2 % n][None, :, None]
              * np.exp(1j * m2 * soft.y)[None, None, :])
    np.testing.assert_allclose(g, expect, atol=1e-12)


def test_average_separation_cubic_lattice_and_kdtree():
    from scipy.spatial import cKDTree
    a = 1.37
    g = np.arange(4) * a
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    assert O.average_separation(X) == pytest.approx(a, rel=1e-14)
    rng = np.random.default_rng(3)
    Y = rng.normal(size=(57, 3))
    d, _ = cKDTree(Y).query(Y, k=2)
    assert O.average_separation(Y) == pytest.approx(d[:, 1].mean(), rel=1e-13)


def _exact_copy_case(seed=13, N=18, L=14, s=0.35):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, 3)) * 1.1
    X -= X.mean(0)
    Rtrue = np.array([0.9, 1.2, 4.0])
    # refine() rotates X2 by X2 . EulerM(R); the exact copy X2 = X1 EulerM(R)^T
    # is therefore mapped back onto X1 at R = Rtrue
    X2 = X.dot(O.euler_matrix(*Rtrue).T)
    I = O.so3_coeffs(X, X2, L, s, fast=True)
    return X, X2, I, Rtrue, L


def test_max_overlap_recovers_exact_rotation():
    # every degree of the overlap of an exact copy peaks at the true rotation
    # (each degree is a positive-definite kernel's inner product), so the
    # local maximiser started at the grid peak must land on it
    X, X2, I, Rtrue, L = _exact_copy_case()
    bw = L + 1
    soft = O.SOFTOracle(bw)
    grid = soft.iSOFT(I)
    ind, find = O.find_max(grid)
    R0 = soft.indtoEuler(find)
    R, res = O.max_overlap(R0, I.conj(), L)
    # L-BFGS-B stops at its default tolerance ~2e-5 rad from the optimum (the
    # grid spacing is 0.2 rad); its objective is then within 1e-8 of the peak
    np.testing.assert_allclose(O.euler_matrix(*R), O.euler_matrix(*Rtrue), atol=1e-4)
    ftrue = O.energy_gradient(Rtrue, I.conj(), L)[0]
    assert ftrue <= res.fun <= ftrue + 1e-8 * abs(ftrue)
    assert res.fun <= O.energy_gradient(R0, I.conj(), L)[0]


def test_max_overlap_end_point_is_a_stationary_local_minimum():
    # a noisy, permuted partner: f(R*) <= f(R0), the gradient vanishes to the
    # optimiser's tolerance, and every axis step of 1e-3 rad raises f
    from fastoverlap_amd import workloads as W
    c = W.config0()
    X1 = c["pos1"] - c["pos1"].mean(0)
    X2 = c["pos2"] - c["pos2"].mean(0)
    L = 15
    I = O.so3_coeffs_species(X1, X2, c["groups"], L, 0.3, fast=True)
    soft = O.SOFTOracle(L + 1)
    _, find = O.find_max(soft.iSOFT(I))
    R0 = soft.indtoEuler(find)
    R, res = O.max_overlap(R0, I.conj(), L)
    f0 = O.energy_gradient(R0, I.conj(), L)[0]
    f, g = O.energy_gradient(R, I.conj(), L)
    assert f <= f0
    assert np.abs(g).max() <= 1e-4 * abs(f)
    for e in np.eye(3):
        for h in (1e-3, -1e-3):
            assert O.energy_gradient(R + h * e, I.conj(), L)[0] > f
    # an independent derivative-free optimiser from the same start agrees
    from scipy.optimize import minimize
    nm = minimize(lambda r: O.energy_gradient(r, I.conj(), L)[0], R0, method="Nelder-Mead",
                  options=dict(xatol=1e-10, fatol=1e-13, maxiter=4000))
    np.testing.assert_allclose(O.euler_matrix(*nm.x), O.euler_matrix(*R), atol=1e-4)
    assert res.fun <= nm.fun + 1e-9 * abs(nm.fun)


def test_periodic_hungarian_brute_force_and_semantics():
    # intended semantics: perm[i] = atom of X2 matched to X1[i] under the
    # minimum-image Euclidean cost, X2[perm] ~ X1, per permutation group
    rng = np.random.default_rng(8)
    box = np.array([3.0, 4.0, 5.0])
    P = O.PeriodicOracle(7, box, [np.arange(4), np.arange(4, 7)], n=2)
    X1 = rng.random((7, 3)) * box
    X2 = rng.random((7, 3)) * box
    dist, perm = P.hungarian(X1, X2)
    cost = P.cost_matrix(X1, X2)
    best = None
    for pa in itertools.permutations(range(4)):
        for pb in itertools.permutations(range(4, 7)):
            p = np.array(pa + pb)
            c = cost[np.arange(7), p].sum()
            if best is None or c < best[0]:
                best = (c, p)
    np.testing.assert_array_equal(perm, best[1])
    d = X1 - X2[perm]
    d -= np.round(d / box) * box
    assert dist == pytest.approx(np.linalg.norm(d), rel=1e-14)
    # a shifted, permuted copy is matched back exactly
    sh = rng.permutation(7)
    sh[:4] = rng.permutation(4)
    sh[4:] = 4 + rng.permutation(3)
    Y = X1[sh] + box * np.array([1, -2, 0])
    _, p2 = P.hungarian(X1, Y)
    np.testing.assert_array_equal(sh[p2], np.arange(7))


# ---- multi-peak search ----------------------------------------------------------------

def test_gaussian_nd_closed_form():
    A, mu = 2.5, 0.3
    s = np.array([1.2, 0.1, -0.2, 0.8, 0.05, 1.5])      # s00 s01 s02 s11 s12 s22
    c = np.array([0.4, -0.3, 0.2])
    x = np.array([[0.0, 1.0, -1.0], [0.5, 0.0, 2.0], [1.0, -1.0, 0.0]])
    d = x - c[:, None]
    q = (s[0] * d[0] ** 2 + s[1] * d[0] * d[1] + s[2] * d[0] * d[2] + s[3] * d[1] ** 2
         + s[4] * d[1] * d[2] + s[5] * d[2] ** 2)
    np.testing.assert_allclose(O.gaussian_nd(x, A, mu, *s, *c), A * np.exp(-q) + mu,
                               rtol=1e-15)


def _gauss_grid(shape, comps, bg=0.0):
    idx = np.indices(shape).reshape(3, -1).astype(float)
    g = np.full(idx.shape[1], bg)
    for A, s, c in comps:
        g += O.gaussian_nd(idx, A, 0.0, *s, *c)
    return g.reshape(shape)


def test_find_peaks_recovers_known_gaussians():
    # three well-separated anisotropic Gaussians: every centre and amplitude
    # is recovered, in order of height (the argmax sequence of the residual)
    comps = [(5.0, [0.9, 0.2, 0.0, 0.7, -0.1, 1.1], [6.3, 7.8, 5.2]),
             (3.0, [0.6, 0.0, 0.1, 0.8, 0.0, 0.5], [17.6, 6.1, 16.4]),
             (1.5, [1.0, -0.2, 0.0, 1.0, 0.1, 0.9], [8.2, 18.7, 17.9])]
    a = _gauss_grid((24, 24, 24), comps, bg=0.25)
    peaks, amp, mean, sigma, f = O.find_peaks(a, npeaks=3, width=2)
    for k, (A, s, c) in enumerate(comps):
        np.testing.assert_allclose(peaks[k], c, atol=1e-6)
        assert amp[k] == pytest.approx(A, rel=1e-6)
    # brute force: the first index of each fit is the argmax of what remains
    assert np.unravel_index(a.argmax(), a.shape) == tuple(np.round(comps[0][2]).astype(int))
    # the residual is flat once every component is removed
    assert np.abs(f).max() < 1e-5 * a.max()


def test_find_peaks_window_wraps_periodically():
    comps = [(4.0, [0.8, 0.0, 0.0, 0.8, 0.0, 0.8], [0.3, 9.6, 12.0])]
    shape = (16, 16, 16)
    idx = np.indices(shape).reshape(3, -1).astype(float)
    # periodic images so the grid itself is periodic around index 0
    g = np.zeros(idx.shape[1])
    for sh in (-16.0, 0.0, 16.0):
        A, s, c = comps[0]
        g += O.gaussian_nd(idx, A, 0.0, *s, *(np.array(c) + [sh, 0, 0]))
    a = g.reshape(shape)
    peaks = O.find_peaks(a, npeaks=1, width=2)[0]
    np.testing.assert_allclose(peaks[0], comps[0][2], atol=1e-6)


def test_malign_exact_rotation_and_agreement_with_align():
    rng = np.random.default_rng(11)
    X = rng.normal(size=(30, 3)) * 1.2
    X -= X.mean(0)
    Y = X.dot(O.euler_matrix(1.0, 0.7, 2.5).T)[rng.permutation(30)]
    assert O.spherical_malign(X, Y, 12, 0.3, nrot=3)[0] < 1e-6
    # an easy pair: the distance depends only on the permutation, which the
    # peak fits and the refined grid maximum agree on
    from fastoverlap_amd import workloads as W
    c = W.config0()
    m = O.spherical_malign(c["pos1"], c["pos2"], 15, 0.3, c["groups"], nrot=4)
    a = O.spherical_align(c["pos1"], c["pos2"], 15, 0.3, perm=c["groups"], fast=True)
    assert m[0] == pytest.approx(a.dist, abs=1e-10)
    np.testing.assert_array_equal(m[3], a.perm)


def test_periodic_find_disps_multi_peak_recovers_translation():
    rng = np.random.default_rng(6)
    box = np.full(3, 6.0)
    X = rng.random((40, 3)) * box
    shift = np.array([1.2, -2.1, 0.45])
    perm = rng.permutation(40)
    Y = (X + shift)[perm]
    P = O.PeriodicOracle(40, box)
    (d1, _, _, p1, disp1), _, _ = P.align(X, Y)
    (d3, _, _, p3, disp3), _, _ = P.align(X, Y, npeaks=3)
    assert d1 < 1e-9 and d3 < 1e-9
    np.testing.assert_array_equal(perm[p3], np.arange(40))


def test_oh_operations_match_the_reference_table():
    """The 48 octahedral operations: the reference's OHOPS table, in its
    order (fixture read from its DATA statement), and a group of distinct
    signed permutation matrices closed under products."""
    ref = json.load(open(os.path.join(GOLDEN, "oh_ops.json")))["ops"]
    M = O.oh_matrices()
    np.testing.assert_array_equal(M, np.array(ref))
    keys = {m.tobytes() for m in M}
    assert len(keys) == 48
    for a in M:
        assert sorted(np.abs(a).sum(0)) == [1, 1, 1] and sorted(np.abs(a).sum(1)) == [1, 1, 1]
        for b in M:
            assert (a @ b).tobytes() in keys
    np.testing.assert_array_equal(M[0], np.eye(3))


def _box_pair(seed, n=32, side=5.0):
    rng = np.random.default_rng(seed)
    box = np.full(3, side)
    X = rng.random((n, 3)) * box
    return rng, box, X


def test_bulk_align_reduces_to_align_and_finds_images():
    rng, box, X = _box_pair(21)
    P = O.PeriodicOracle(32, box)
    Y = X + rng.normal(scale=0.05, size=X.shape)
    # one displacement, no Oh search: the Python drop-in's align
    (d, _, X2, p, _), o = O.bulk_align(P, X, Y)
    (da, _, X2a, pa, _), _, _ = P.align(X, Y)
    assert o == 0 and d == da
    np.testing.assert_array_equal(p, pa)
    # a translated, permuted copy: operation 0 (identity) wins at distance 0
    perm = rng.permutation(32)
    Z = (X + np.array([0.9, -1.7, 2.2]))[perm]
    (d1, _, _, p1, _), o1 = O.bulk_align(P, X, Z, ohcell=1)
    assert d1 < 1e-9 and o1 == 0
    np.testing.assert_array_equal(perm[p1], np.arange(32))
    # an octahedral image: with per-image overlaps the inverse operation
    # brings it back to a translated copy (distance 0), and no earlier
    # operation does
    M = O.oh_matrices()
    k = 13
    W = (X @ M[k].T + np.array([0.4, 1.3, -0.8]))[perm]
    (d2, _, _, p2, _), o2 = O.bulk_align(P, X, W, ohcell=2)
    np.testing.assert_array_equal(M[o2] @ M[k], np.eye(3))
    assert d2 < 1e-9
    np.testing.assert_array_equal(perm[p2], np.arange(32))
    # several displacements: each refined, the best kept (never worse than
    # the single interpolated maximum)
    (d3, _, _, _, _), _ = O.bulk_align(P, X, Y, ndisps=4)
    assert d3 <= d + 1e-12


def test_bulk_defaults_closed_form():
    kw, nwave, nf = O.bulk_defaults(256, np.full(3, 5.975206329))
    assert kw == pytest.approx(5.975206329 / 256 ** (1 / 3) / 3, rel=1e-15)
    # 2 pi / L * 1.5 / kw = 2 pi * 4.5 * 256^(1/3) / L^2 = 5.03  -> 6; FFT
    # side: the 2,3,5-smooth length >= 4 * 6 + 3 = 27
    assert 2 * math.pi * 4.5 * 256 ** (1 / 3) / 5.975206329 ** 2 == pytest.approx(5.03, abs=0.01)
    assert nwave == 6 and nf == 27


def test_bulk_align_group_cells_are_pair_alignments():
    rng, box, X = _box_pair(22, n=24, side=4.5)
    coords = np.stack([X, (X + 0.3)[rng.permutation(24)], rng.random((24, 3)) * box])
    dist, aligned = O.bulk_align_group(coords, boxvec=box)
    kw, nwave, nf = O.bulk_defaults(24, box)
    P = O.PeriodicOracle(24, box, scale=kw, n=nwave, fshape=(nf,) * 3)
    for j in range(3):
        for i in range(3):
            if i < j:
                assert dist[i, j] == 0 and not aligned[i, j].any()
                continue
            (d, _, X2, _, _), _ = O.bulk_align(P, coords[i], coords[j])
            assert dist[i, j] == d
            np.testing.assert_array_equal(aligned[i, j], X2)
    assert dist[1, 0] < 1e-9 and dist[0, 0] < 1e-9


def test_harmonic_basis_is_orthonormal_and_laguerre():
    from scipy.integrate import quad
    from scipy.special import gamma
    for l in (0, 3):
        phis = []
        for n in range(5):
            g = O.harmonic_coeffs(n, l, 1.0)
            phis.append(lambda r, g=g: np.polyval(g[::-1], r) * np.exp(-0.5 * r * r))
            # Laguerre closed form of the coefficients at s = l + 2k
            a = l + 0.5
            for k in range(n + 1):
                v = (-1) ** k * gamma(n + a + 1) / (gamma(n - k + 1) * gamma(a + k + 1)
                                                    * math.factorial(k))
                assert g[l + 2 * k] == pytest.approx(v * O.harmonic_norm(n, l, 1.0), rel=1e-10)
        for i in range(5):
            for j in range(5):
                v = quad(lambda r: phis[i](r) * phis[j](r) * r * r, 0, 20, limit=200)[0]
                assert v == pytest.approx(1.0 if i == j else 0.0, abs=1e-9)
    np.testing.assert_array_equal(O.harmonic_coeff_table(3, 2, 1.3)[2, 1, :6],
                                  O.harmonic_coeffs(2, 1, 1.3))


def test_harmonic_radial_integral_matches_quadrature():
    from scipy.integrate import quad
    from scipy.special import spherical_in
    for (s, l, rj, sig, r0) in [(0, 0, 1.3, 0.4, 1.1), (4, 2, 0.7, 0.3, 0.8),
                                (9, 5, 2.2, 0.35, 1.0), (6, 6, 0.0, 0.3, 1.2)]:
        f = (lambda r: np.exp(-(r * r + rj * rj) / (2 * sig ** 2)) * np.exp(-r * r / (2 * r0 ** 2))
             * spherical_in(l, r * rj / sig ** 2) * r ** s * r * r)
        v = quad(f, 0, rj + 12 * sig, limit=400, epsabs=0, epsrel=1e-12)[0]
        # the closed form carries the 4 pi of the Gaussian's multipole
        # expansion exp(-|x - r_j|^2 / 2 s^2) = 4 pi e^{..} sum_l i_l Y Y*
        w = O.radial_integral_harmonic(s, l, rj, sig, r0) / (4 * math.pi)
        if l > 0 and rj == 0:
            assert w == 0 and v == 0
        else:
            assert w == pytest.approx(v, rel=1e-9)


def test_harmonic_coefficients_layout_and_rotation_invariance():
    rng = np.random.default_rng(41)
    J, nmax, sig, r0 = 6, 4, 0.35, 1.2
    # one atom on the z axis: only m = 0 survives, Y_l0 = sqrt((2l+1)/4pi)
    c = O.harmonic_coeff(np.array([[0.0, 0.0, 0.9]]), J, nmax, sig, r0)
    tab = O.harmonic_coeff_table(nmax, J, r0)
    for l in range(J + 1):
        for n in range(nmax + 1):
            d = sum(tab[n, l, s] * O.radial_integral_harmonic(s, l, 0.9, sig, r0)
                    for s in range(2 * nmax + J + 1))
            assert c[n, l, 0] == pytest.approx(d * math.sqrt((2 * l + 1) / (4 * math.pi)), rel=1e-12)
        m = np.arange(2 * J + 1) != 0
        assert np.abs(c[:, l, m]).max() <= 1e-15 * np.abs(c).max()
    # the power per (n, l) is invariant under rotations
    X = rng.normal(size=(9, 3))
    Rm = O.euler_matrix(0.4, 1.1, 2.0)
    p1 = (np.abs(O.harmonic_coeff(X, J, nmax, sig, r0)) ** 2).sum(2)
    p2 = (np.abs(O.harmonic_coeff(X @ Rm, J, nmax, sig, r0)) ** 2).sum(2)
    np.testing.assert_allclose(p1, p2, rtol=1e-10, atol=1e-14 * p1.max())
    # a structure with itself: Hermitian per degree; -X2 flips odd degrees
    I = O.harmonic_so3_coeffs(X, X, [np.arange(9)], J, nmax, sig, r0)
    np.testing.assert_allclose(I, I.transpose(0, 2, 1).conj(), rtol=0, atol=1e-14 * np.abs(I).max())
    Y = rng.normal(size=(9, 3))
    Ia = O.harmonic_so3_coeffs(X, Y, [np.arange(9)], J, nmax, sig, r0)
    Ib = O.harmonic_so3_coeffs(X, -Y, [np.arange(9)], J, nmax, sig, r0)
    sg = (-1.0) ** np.arange(J + 1)
    np.testing.assert_allclose(Ib, Ia * sg[:, None, None], rtol=0, atol=1e-13 * np.abs(Ia).max())


def test_harmonic_align_exact_copies():
    rng = np.random.default_rng(42)
    X = rng.normal(size=(14, 3)) * 1.1
    perm = rng.permutation(14)
    Rm = O.euler_matrix(2.1, 0.9, 0.3)
    r = O.harmonic_align(X, (X @ Rm.T)[perm], 8, 5, scale=0.3, r0=1.0)
    assert r.dist < 1e-6
    np.testing.assert_array_equal(perm[r.perm], np.arange(14))
    r = O.harmonic_align(X, -(X @ Rm.T)[perm], 8, 5, scale=0.3, r0=1.0)
    assert r.dist < 1e-6 and r.inverted


def test_fortran_harmonic_recurrence_matches_quadrature():
    from scipy.integrate import quad
    from scipy.special import eval_genlaguerre, gamma, spherical_in
    N, L, rj, sig, r0 = 4, 6, 1.1, 0.35, 1.3
    R = O.harmonic_nl(N, L + 2 * N, rj, sig, r0)

    def g(n, l, r):
        nn = math.sqrt(2 * math.factorial(n) / (r0 ** 3 * gamma(n + l + 1.5)))
        return nn * (r / r0) ** l * np.exp(-r * r / (2 * r0 * r0)) * \
            eval_genlaguerre(n, l + 0.5, (r / r0) ** 2)
    for n in range(N + 1):
        # the oscillator functions are orthonormal under r^2 dr
        assert quad(lambda r: g(n, 2, r) ** 2 * r * r, 0, 20)[0] == pytest.approx(1, abs=1e-10)
        for l in range(0, L + 1, 2):
            v = 4 * math.pi * quad(lambda r: g(n, l, r) * np.exp(-(r * r + rj * rj) / (2 * sig ** 2))
                                   * spherical_in(l, r * rj / sig ** 2) * r * r, 0, 15,
                                   limit=300, epsabs=0, epsrel=1e-12)[0]
            assert R[n, l] == pytest.approx(v, rel=1e-10)


def test_fortran_harmonic_coefficients_and_similarity():
    rng = np.random.default_rng(61)
    J, nmax, hw, kw = 6, 4, 1.2, 0.35
    c = O.harmonic_coeff_fortran(np.array([[0.0, 0.0, 0.8]]), J, nmax, hw, kw)
    R = O.harmonic_nl(nmax, J + 2 * nmax, 0.8, kw, hw)
    for l in range(J + 1):
        np.testing.assert_allclose(c[:, l, 0], R[:, l] * math.sqrt((2 * l + 1) / (4 * math.pi)),
                                   rtol=1e-12)
    X = rng.normal(size=(10, 3))
    X -= X.mean(0)
    Rm = O.euler_matrix(0.3, 1.2, 2.2)
    cX = O.harmonic_coeff_fortran(X, J, nmax, hw, kw)
    cR = O.harmonic_coeff_fortran(X @ Rm.T, J, nmax, hw, kw)
    np.testing.assert_allclose((np.abs(cX) ** 2).sum(2), (np.abs(cR) ** 2).sum(2), rtol=1e-10)
    # the norm is unchanged by rotating the partner; the maximum is the grid's
    soft = O.SOFTOracle(J + 1)
    n_self, m_self = O.harmonic_similarity([cX], [cX], J, soft)
    n_rot, m_rot = O.harmonic_similarity([cX], [cR], J, soft)
    assert n_rot == pytest.approx(n_self, rel=1e-10)
    I = np.einsum("nlm,nlo->lmo", cX.conj(), cR)
    assert m_rot == soft.iSOFT(I).real.max()
    avg, mx, navg, nmx = O.harmonic_compare_list([X, X @ Rm.T, rng.normal(size=(10, 3))], J,
                                                 nmax, hw, kw)
    np.testing.assert_array_equal(avg, avg.T)
    np.testing.assert_allclose(np.diag(navg), 1.0, rtol=1e-15)
    np.testing.assert_allclose(np.diag(nmx), 1.0, rtol=1e-15)
    assert avg[0, 1] == pytest.approx(avg[0, 0], rel=1e-10)
