"""``PeriodicAlignFortran`` (the bulk Fortran entry points: ALIGN with
NDISPLACEMENTS and OHCELLT, ALIGNGROUP) against the FP64 oracle's
restatement of ``fastoverlap/f90/fastbulk.f90:336-565``: distance and
aligned coordinates within 1e-8, permutation identical."""

import numpy as np
import pytest

import fastoverlap_amd as fo
from oracle import fo_oracle as O

pytestmark = pytest.mark.gpu

BOX = np.ones(3) * 5.975206329
GROUPS = [np.arange(204), np.arange(204, 256)]


def _oracle(n=256, groups=GROUPS, box=BOX):
    return O.PeriodicOracle(n, box, groups)


def _check(res, ora):
    dist, _, X2, perm = res
    (odist, _, oX2, operm, _), _ = ora
    assert abs(dist - odist) <= 1e-8
    np.testing.assert_array_equal(perm - 1, operm)
    np.testing.assert_allclose(X2, oX2, rtol=0, atol=1e-8)


def test_oh_operation_table_matches_oracle(gpu):
    L = fo._lib.load()
    m = np.zeros(9)
    for o in range(48):
        assert L.fo_oh_operation(o, fo._lib.dptr(m)) == 0
        np.testing.assert_array_equal(m.reshape(3, 3), O.oh_matrices()[o])


@pytest.mark.parametrize("ohcell", [0, 1, 2])
def test_bulk_align_ohcell_matches_oracle(gpu, blj256, ohcell):
    print(f"case: BLJ256 bulk align, ohcell={ohcell}", flush=True)
    q1, q2 = blj256
    A = fo.PeriodicAlignFortran(256, BOX, perm=GROUPS)
    _check(A.align(q1, q2, ndisps=1, ohcell=ohcell),
           O.bulk_align(_oracle(), q1, q2, ohcell=ohcell))


def test_octahedral_image_is_recovered(gpu):
    """An octahedral image of a structure, translated and permuted: with
    per-image overlaps the search brings it back to distance 0."""
    print("case: 96-atom octahedral image, ohcell=2", flush=True)
    rng = np.random.default_rng(31)
    box = np.full(3, 5.0)
    X = rng.random((96, 3)) * box
    perm = rng.permutation(96)
    W = (X @ O.oh_matrices()[29].T + np.array([0.7, -1.1, 2.3]))[perm]
    A = fo.PeriodicAlignFortran(96, box)
    res = A.align(X, W, ndisps=1, ohcell=2)
    assert res[0] < 1e-8
    np.testing.assert_array_equal(perm[res[3] - 1], np.arange(96))
    _check(res, O.bulk_align(O.PeriodicOracle(96, box), X, W, ohcell=2))


def test_several_displacements_match_oracle(gpu):
    print("case: 64-atom noisy translated copy, ndisps=4", flush=True)
    rng = np.random.default_rng(32)
    box = np.full(3, 4.5)
    X = rng.random((64, 3)) * box
    Y = (X + np.array([1.3, 0.4, -2.0]) + rng.normal(scale=0.04, size=X.shape))[
        rng.permutation(64)]
    A = fo.PeriodicAlignFortran(64, box)
    _check(A.align(X, Y, ndisps=4), O.bulk_align(O.PeriodicOracle(64, box), X, Y, ndisps=4))


def test_align_group_matches_oracle(gpu, blj256):
    print("case: ALIGNGROUP of four BLJ256 frames", flush=True)
    q1, q2 = blj256
    rng = np.random.default_rng(33)
    frames = np.stack([q1, q2, (q1 + 0.8)[np.r_[rng.permutation(204), 204 + rng.permutation(52)]],
                       q2 * np.array([1.0, -1.0, 1.0])])
    A = fo.PeriodicAlignFortran(256, BOX, perm=GROUPS)
    dists, aligned = A.alignGroup(frames, ndisps=1)
    odist, oal = O.bulk_align_group(frames, boxvec=BOX, perm=GROUPS)
    np.testing.assert_allclose(dists, odist, rtol=0, atol=1e-8)
    np.testing.assert_allclose(aligned, oal.transpose(2, 3, 0, 1), rtol=0, atol=1e-8)
    assert dists[2, 0] < 1e-8 and dists[0, 0] < 1e-8
