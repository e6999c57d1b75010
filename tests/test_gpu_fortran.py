"""The Fortran caller (iso_c_binding modules, built in-tree by
fastoverlap_amd/build.py) against the FP64 oracle: distance within 1e-8 and
the 1-based BESTPERM equal to the oracle's permutation, for the cluster
(LJ38) and the bulk (BLJ256) ALIGN, the bulk ALIGN with OHCELLT and the group
alignment ALIGNGROUP (with CALCDEFAULTS), and the harmonic-basis ALIGNHARM
and CALCOVERLAPMATRICES."""

import os
import subprocess
import sys

import numpy as np
import pytest

import fastoverlap_amd as fo
from oracle import fo_oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fastoverlap_amd", "fortran", "_build", "test_align")


def test_fortran_align_matches_dropin(gpu, lj38, blj256):
    assert os.path.exists(EXE), "build the Fortran caller with python -m fastoverlap_amd.build"
    print("case: Fortran test_align LJ38 (L=14) and BLJ256", flush=True)
    r = subprocess.run([EXE, os.path.join(ROOT, "tests", "golden")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    cdist, cstat = float(out["CLUSTER"][0]), int(out["CLUSTER"][1])
    bdist, bstat = float(out["BULK"][0]), int(out["BULK"][1])
    assert cstat == 0 and bstat == 0
    # second calls with the same system reuse the first call's batch
    assert out["CLUSTER2"] == out["CLUSTER"] and out["CPERM2"] == out["CPERM"]
    assert out["BULK2"] == out["BULK"]
    p1, p2 = lj38
    o = O.spherical_align(p1, p2, 14, 0.3, fast=True)
    assert abs(cdist - o.dist) <= 1e-8
    cperm = np.array([int(v) for v in out["CPERM"]]) - 1
    np.testing.assert_array_equal(cperm, o.perm)
    q1, q2 = blj256
    box = np.ones(3) * 5.975206329
    groups = [np.arange(204), np.arange(204, 256)]
    (odist, _, _, operm, _), _, _ = O.PeriodicOracle(256, box, groups).align(q1, q2)
    assert bdist == pytest.approx(1.559, abs=1e-3)
    assert abs(bdist - odist) <= 1e-8
    bperm = np.array([int(v) for v in out["BPERM"]]) - 1
    np.testing.assert_array_equal(bperm, operm)
    # and the Python drop-in agrees with both
    assert abs(fo.PeriodicAlign(256, box, groups)(q1, q2)[0] - bdist) <= 1e-8
    # OHCELLT: 48 images from the displacement of the untransformed pair
    assert int(out["BULKOH"][1]) == 0
    (odist, _, _, operm, _), _ = O.bulk_align(O.PeriodicOracle(256, box, groups), q1, q2,
                                              ohcell=1)
    assert abs(float(out["BULKOH"][0]) - odist) <= 1e-8
    np.testing.assert_array_equal(np.array([int(v) for v in out["OHPERM"]]) - 1, operm)
    # ALIGNGROUP over (coords, finish, coords mirrored in z), symmetric
    kw, nwave, nf = O.bulk_defaults(256, box)
    assert float(out["DEFAULTS"][0]) == pytest.approx(kw, rel=1e-14)
    assert [int(v) for v in out["DEFAULTS"][1:]] == [nwave, nf, 0]
    frames = np.stack([q1, q2, q1 * np.array([1.0, 1.0, -1.0])])
    odm, _ = O.bulk_align_group(frames, boxvec=box, perm=groups)
    dm = np.array([float(v) for v in out["GROUP"]]).reshape(3, 3, order="F")
    np.testing.assert_allclose(dm, odm, rtol=0, atol=1e-8)
    # ALIGNHARM: the Fortran module's harmonic basis (degree 8, scale 1.0)
    assert int(out["HARM"][1]) == 0
    oh = O.harmonic_align(p1, p2, 14, 8, scale=0.3, r0=1.0, basis="fortran")
    assert abs(float(out["HARM"][0]) - oh.dist) <= 1e-8
    np.testing.assert_array_equal(np.array([int(v) for v in out["HPERM"]]) - 1, oh.perm)
    # CALCOVERLAPMATRICES of (coords, finish, coords), centred
    frames = [p1 - p1.mean(0), p2 - p2.mean(0), p1 - p1.mean(0)]
    avg, mx, _, _ = O.harmonic_compare_list(frames, 10, 8, 1.0, 0.3)
    norms = np.array([float(v) for v in out["NORMS"]]).reshape(3, 3, order="F")
    maxov = np.array([float(v) for v in out["MAXOV"]]).reshape(3, 3, order="F")
    np.testing.assert_allclose(norms, avg, rtol=1e-10)
    np.testing.assert_allclose(maxov, mx, rtol=1e-10)
