"""``SphericalHarmonicAlign`` on the device (harmonic-oscillator basis
coefficients, fastoverlap/sphericalAlignment.py:276-414) against the FP64
oracle's restatement: per-structure coefficients and C^l within 1e-10 of the
largest entry, argmax of both orientations and permutation identical,
distance within 1e-8."""

import numpy as np
import pytest

import fastoverlap_amd as fo
from fastoverlap_amd import workloads as W
from oracle import fo_oracle as O

pytestmark = pytest.mark.gpu


def _cloud(seed, n, radius=1.6):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 3))
    X *= radius / np.abs(X).max()
    return rng, X - X.mean(0)


def test_harmonic_coefficients_match_oracle(gpu):
    print("case: 20-atom cloud, Jmax 10, nmax 6, harmscale 1.1", flush=True)
    _, X = _cloud(51, 20)
    A = fo.SphericalHarmonicAlign(scale=0.35, harmscale=1.1, nmax=6, Jmax=10)
    c = A.calcHarmCoeff(X)
    o = O.harmonic_coeff(X, 10, 6, 0.35, 1.1)
    np.testing.assert_allclose(c, o, rtol=0, atol=1e-10 * np.abs(o).max())
    np.testing.assert_allclose(A.coeffs, O.harmonic_coeff_table(6, 10, 1.1), rtol=1e-13,
                               atol=0)
    _, Y = _cloud(52, 20)
    I = A.calcSO3Coeffs(X, Y)
    Io = O.harmonic_so3_coeffs(X, Y, [np.arange(20)], 10, 6, 0.35, 1.1)
    np.testing.assert_allclose(I, Io, rtol=0, atol=1e-10 * np.abs(Io).max())


def _check(A, p1, p2, groups, Jmax, nmax, scale, r0):
    d, X1, X2 = A.align(p1, p2)
    o = O.harmonic_align(p1, p2, Jmax, nmax, scale=scale, r0=r0,
                         perm=None if groups is None else list(groups))
    pk = A.last["peaks"]
    for k in range(2):
        assert tuple(int(v) for v in pk["ind"][0, k]) == o.ind[k], (k, pk["ind"][0, k], o.ind[k])
    assert A.last["inverted"] == o.inverted
    np.testing.assert_array_equal(A.last["perm"], o.perm)
    assert abs(d - o.dist) <= 1e-8
    np.testing.assert_allclose(X2, o.X2, rtol=0, atol=1e-7)


def test_harmonic_align_lj38_matches_oracle(gpu, lj38):
    print("case: LJ38 harmonic align, Jmax 15, nmax 15 (reference defaults)", flush=True)
    p1, p2 = lj38
    A = fo.SphericalHarmonicAlign(scale=0.3, harmscale=1.0, nmax=15, Jmax=15)
    _check(A, p1, p2, None, 15, 15, 0.3, 1.0)


def test_harmonic_align_two_species_matches_oracle(gpu):
    print("case: config-0 style two-species pair, automatic width, harmscale 1.3", flush=True)
    X = W.config0(3)["pos1"]
    rng = np.random.default_rng(53)
    n = len(X)
    groups = [np.arange(0, 26), np.arange(26, n)]
    # partner: rotated, permuted within species, N(0, 0.02) noise
    pm = np.concatenate([rng.permutation(26), 26 + rng.permutation(n - 26)])
    Y = (X @ O.euler_matrix(1.9, 0.6, 4.0).T)[pm] + rng.normal(scale=0.02, size=X.shape)
    A = fo.SphericalHarmonicAlign(harmscale=1.3, nmax=8, Jmax=12, perm=groups)
    A.align(X, Y)
    scale = A.scale
    assert scale == pytest.approx((O.average_separation(X) + O.average_separation(Y)) / 6,
                                  rel=1e-12)
    _check(A, X, Y, groups, 12, 8, scale, 1.3)


def test_harmonic_exact_copies_and_group(gpu):
    print("case: exact rotated and inverted copies, alignGroup of three", flush=True)
    rng, X = _cloud(54, 24)
    Rm = O.euler_matrix(0.8, 2.0, 1.4)
    perm = rng.permutation(24)
    Y = (X @ Rm.T)[perm]
    A = fo.SphericalHarmonicAlign(scale=0.3, nmax=10, Jmax=12)
    assert A.align(X, Y)[0] < 1e-6
    assert A.align(X, -Y)[0] < 1e-6 and A.last["inverted"]
    Z = _cloud(55, 24)[1]
    coords = [X, Y, Z]
    dists, aligned = A.alignGroup(coords, keepCoords=True)
    for i in range(3):
        for j in range(3):
            o = O.harmonic_align(coords[i], coords[j], 12, 10, scale=0.3, r0=1.0)
            assert abs(dists[i, j] - o.dist) <= 1e-8
            np.testing.assert_allclose(aligned[1, i, j], o.X2, rtol=0, atol=1e-7)
    assert dists[0, 1] < 1e-6 and dists[1, 0] < 1e-6


def test_fortran_basis_coefficients_match_oracle(gpu):
    print("case: Fortran HARMONICCOEFFS basis, 18 atoms, Jmax 9, nmax 7", flush=True)
    _, X = _cloud(56, 18)
    with fo.SphericalBatch([18], [9], [0.32], 20, None,
                           fo._lib.SPH_NO_INVERT | fo._lib.SPH_NO_CENTER,
                           harmonic=(7, 1.25, 1)) as b:
        b.set_positions(X, X)
        b.run(fo._lib.STAGE_PREP | fo._lib.STAGE_COEFF)
        c, _ = b.harmonic_coeffs(0)
    o = O.harmonic_coeff_fortran(X, 9, 7, 1.25, 0.32)
    np.testing.assert_allclose(c[0], o, rtol=0, atol=1e-10 * np.abs(o).max())


def test_compare_list_matches_oracle(gpu):
    print("case: compareList of five 20-atom structures, Jmax 8, nmax 6", flush=True)
    rng = np.random.default_rng(57)
    base = [_cloud(60 + k, 20)[1] for k in range(3)]
    Rm = O.euler_matrix(1.0, 0.5, 2.5)
    coords = [base[0], base[0] @ Rm.T, base[1], base[2], base[1][rng.permutation(20)] + 0.4]
    A = fo.SphericalHarmonicAlignFortran(scale=0.3, Jmax=8, harmscale=1.1, nmax=6)
    got = A.compareList(coords)
    exp = O.harmonic_compare_list(coords, 8, 6, 1.1, 0.3)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, rtol=1e-10, atol=0)
    # a rotated copy has the self-overlap norm; a shifted, permuted copy too
    assert got[0][0, 1] == pytest.approx(got[0][0, 0], rel=1e-10)
    assert got[2][2, 4] == pytest.approx(1.0, rel=1e-10)


def test_fortran_harmonic_align_matches_oracle(gpu, lj38):
    print("case: LJ38 through SphericalHarmonicAlignFortran (Jmax 15, nmax 20)", flush=True)
    p1, p2 = lj38
    A = fo.SphericalHarmonicAlignFortran(scale=0.3, Jmax=15, harmscale=1.0, nmax=20)
    d, X1, X2, rmat = A(p1, p2)
    o = O.harmonic_align(p1, p2, 15, 20, scale=0.3, r0=1.0, basis="fortran")
    assert abs(d - o.dist) <= 1e-8
    np.testing.assert_array_equal(A.last["perm"], o.perm)
    assert A.last["inverted"] == o.inverted
    X2c = (p2 - p2.mean(0)) * (-1.0 if o.inverted else 1.0)
    np.testing.assert_allclose(X2, (X2c @ rmat.T)[o.perm], rtol=0, atol=1e-8)
