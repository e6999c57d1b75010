"""GPU parity of the periodic path against the FP64 oracle (identical
argmax and permutation, distance within 1e-8, |f| grid within 1e-5 of its
maximum)."""

import sys

import numpy as np
import pytest

import fastoverlap_amd as fo
from fastoverlap_amd import _lib, workloads as W
from fastoverlap_amd.batch import PeriodicBatch
from oracle import fo_oracle as O

pytestmark = pytest.mark.gpu

DIST_TOL = 1e-8
GRID_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_structure_factors_match_oracle(gpu):
    say("case: calcFourierCoeff N=40 n=5")
    rng = np.random.default_rng(0)
    box = np.array([5.0, 6.0, 5.5])
    X = rng.random((40, 3)) * box * 1.7 - 2.0   # unwrapped coordinates
    groups = [np.arange(30), np.arange(30, 40)]
    pa = fo.PeriodicAlign(40, box, groups, n=5)
    C = pa.calcFourierCoeff(X)
    ref = O.PeriodicOracle(40, box, groups, n=5).calc_fourier_coeff(X)
    np.testing.assert_allclose(C, ref, atol=1e-11 * np.abs(ref).max())


def test_blj256_known_answer_dropin(gpu, blj256):
    say("case: BLJ256 PeriodicAlign drop-in (n=9, F=40)")
    p1, p2 = blj256
    box = np.ones(3) * 5.975206329
    groups = [np.arange(204), np.arange(204, 256)]
    pa = fo.PeriodicAlign(256, box, groups)
    dist, X1, X2, perm, disp = pa(p1, p2)
    P = O.PeriodicOracle(256, box, groups)
    (odist, oX1, oX2, operm, odisp), oind, ofabs = P.align(p1, p2)
    assert dist == pytest.approx(1.559, abs=1e-3)
    assert abs(dist - odist) <= DIST_TOL
    np.testing.assert_array_equal(perm, operm)
    assert tuple(pa.last["peaks"]["ind"][0]) == oind
    np.testing.assert_allclose(X2, oX2, atol=1e-9)
    # the attributes after a plain call describe that call's configuration:
    # |f|, the complex transform f (phase included) and the product C
    assert np.abs(pa.fabs - ofabs).max() <= GRID_TOL * ofabs.max()
    C, _ = P.overlap_grid(p1, p2)
    f = np.fft.fftn(C, P.fshape, axes=(0, 1, 2))
    assert np.abs(pa.f - f).max() <= GRID_TOL * np.abs(f).max()
    np.testing.assert_allclose(pa.C, C, atol=1e-11 * np.abs(C).max())
    np.testing.assert_allclose(np.abs(pa.f), pa.fabs, rtol=1e-5, atol=1e-6 * ofabs.max())


def test_c1_parity_64_grid(gpu):
    c = W.config1()
    say("case: C1 N=256 nk=8 F=64")
    with PeriodicBatch([256], c["box"], [0.3], 8, 64, [c["groups"]], 10,
                       _lib.PER_KEEP_GRID) as b:
        b.set_positions(c["pos1"], c["pos2"])
        b.run()
        r = b.results()
        pk = b.peaks()
        grid = b.grid(0)
    P = O.PeriodicOracle(256, c["box"], c["groups"], scale=0.3, n=8, fshape=(64, 64, 64))
    (odist, oX1, oX2, operm, odisp), oind, ofabs = P.align(c["pos1"], c["pos2"])
    assert tuple(pk["ind"][0]) == oind
    np.testing.assert_array_equal(r["perm"][0], operm)
    assert abs(r["dist"][0] - odist) <= DIST_TOL
    assert np.abs(grid - ofabs).max() <= GRID_TOL * ofabs.max()
    # the partner is a noisy permuted translate: the LAP recovers it
    np.testing.assert_array_equal(c["true_perm"][r["perm"][0]], np.arange(256))


def test_quick_align_with_coefficients(gpu, blj256):
    say("case: BLJ256 align with precomputed coefficients (Cs=)")
    p1, p2 = blj256
    box = np.ones(3) * 5.975206329
    groups = [np.arange(204), np.arange(204, 256)]
    pa = fo.PeriodicAlign(256, box, groups)
    c1, c2 = pa.calcFourierCoeff(p1), pa.calcFourierCoeff(p2)
    a = pa.align(p1, p2)
    b = pa.align(p1, p2, [c1, c2])
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[3], b[3])


def test_align_group_matches_pairwise(gpu):
    say("case: alignGroup 3 structures N=64")
    rng = np.random.default_rng(8)
    box = np.full(3, 4.2)
    base = rng.random((64, 3)) * box
    coords = [base + rng.normal(scale=0.05, size=base.shape) + s
              for s in (0.0, 1.3, -0.6)]
    pa = fo.PeriodicAlign(64, box)
    D = pa.alignGroup(coords)
    P = O.PeriodicOracle(64, box)
    for i in range(3):
        for j in range(3):
            (od, *_), _, _ = P.align(coords[i], coords[j])
            assert abs(D[i, j] - od) <= DIST_TOL


def test_periodic_batch_independence(gpu):
    say("case: periodic batch of 6 vs alone")
    rng = np.random.default_rng(3)
    box = np.full(3, 5.0)
    p1, p2 = [], []
    for k in range(6):
        X = rng.random((90, 3)) * box
        p1.append(X)
        p2.append(X[rng.permutation(90)] + rng.random(3) * box + rng.normal(scale=0.02,
                                                                           size=X.shape))
    s, nk, F = fo.periodic_defaults(90, box)
    with PeriodicBatch([90] * 6, box, s, nk, F) as b:
        b.set_positions(p1, p2)
        b.run()
        r = b.results()
    with PeriodicBatch([90], box, s, nk, F) as b1:
        b1.set_positions(p1[4], p2[4])
        b1.run()
        r1 = b1.results()
    assert r1["dist"][0] == r["dist"][4]
    np.testing.assert_array_equal(r1["perm"][0], r["perm"][4])


def test_dropin_repeated_calls_reuse_the_batch(gpu):
    """Repeated PeriodicAlign calls on one instance reuse its device batch
    (graph replays after the first call); every result is identical to a
    fresh instance's."""
    c = W.config1()
    partners = [c["pos2"], c["pos2"] + np.array([0.7, -1.3, 2.2]), c["pos2"]]
    pa = fo.PeriodicAlign(256, c["box"], c["groups"], scale=0.3, n=8)
    for k, p2 in enumerate(partners):
        say(f"case: C1 partner {k} on one instance")
        d, X1, X2, perm, disp = pa.align(c["pos1"], p2)
        fresh = fo.PeriodicAlign(256, c["box"], c["groups"], scale=0.3, n=8)
        df, X1f, X2f, permf, dispf = fresh.align(c["pos1"], p2)
        assert d == df, (k, d, df)
        np.testing.assert_array_equal(perm, permf)
        np.testing.assert_array_equal(disp, dispf)
        np.testing.assert_array_equal(X2, X2f)
