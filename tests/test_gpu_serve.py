"""The continuous-batching server: every result is the same alone and in the
mix.

A mixed stream (spherical clusters of uneven N, Jmax and species counts,
periodic boxes of uneven N and transform size) is submitted at once, so that
requests are batched together and batches of different classes overlap on
several streams.  Each request is then served alone (batch of one).  Both
must agree exactly: every per-pair quantity is computed by the pair's own
workgroups with fixed reduction orders, so batching changes nothing (the
stated tolerance is 0 for the permutation and the argmax and 1e-12 for the
distance).  A few small spherical requests are also checked against the FP64
oracle (distance within 1e-8, identical permutation).
"""

import sys

import numpy as np
import pytest

from fastoverlap_amd import workloads as W
from fastoverlap_amd.serve import Server
from oracle import fo_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _requests():
    reqs = [W.config4(p) for p in range(14)]
    # extra small clusters of the batched mix (cheap to check with the oracle)
    for p in range(3):
        c = W.config2(p)
        reqs.append(dict(kind=0, pos1=c["pos1"], pos2=c["pos2"], groups=c["groups"], L=c["L"],
                         grid_n=c["grid_n"], scale=c["scale"]))
    return reqs


def test_server_alone_equals_mix(gpu):
    reqs = _requests()
    kinds = [r["kind"] for r in reqs]
    say(f"case: server mix of {len(reqs)} requests "
        f"({kinds.count(0)} spherical, {kinds.count(1)} periodic), 3 streams")
    with Server(gpu, max_batch=64, streams=3, max_wait_ms=5.0) as srv:
        ids = [srv.submit(r) for r in reqs]
        mixed = [srv.wait(i, timeout_ms=120000.0) for i in ids]
        say("case: the same requests alone")
        alone = [srv.align(r) for r in reqs]
        st = srv.stats()
    assert st["completed"] == 2 * len(reqs)
    assert max(m["batch_size"] for m in mixed) > 1, "the mix was never batched"
    for k, (m, a, r) in enumerate(zip(mixed, alone, reqs)):
        tag = f"request {k} kind {r['kind']} N={len(r['pos1'])}"
        assert m["status"] == 0 and a["status"] == 0, tag
        assert a["batch_size"] == 1, tag
        assert abs(m["dist"] - a["dist"]) <= 1e-12, (tag, m["dist"], a["dist"])
        np.testing.assert_array_equal(m["perm"], a["perm"], err_msg=tag)
        np.testing.assert_array_equal(m["ind"], a["ind"], err_msg=tag)
        assert m["inverted"] == a["inverted"], tag
        np.testing.assert_array_equal(m["X2"], a["X2"], err_msg=tag)
    # oracle parity of the small spherical requests
    for k in range(len(reqs) - 3, len(reqs)):
        r = reqs[k]
        say(f"case: oracle check of request {k} N={len(r['pos1'])} L={r['L']}")
        o = O.spherical_align(r["pos1"], r["pos2"], r["L"], r["scale"], perm=r["groups"],
                              bw=r["grid_n"] // 2, fast=True)
        assert abs(mixed[k]["dist"] - o.dist) <= 1e-8, (k, mixed[k]["dist"], o.dist)
        np.testing.assert_array_equal(mixed[k]["perm"], o.perm)
        assert mixed[k]["inverted"] == o.inverted


def test_server_rejects_bad_request_without_stopping(gpu):
    say("case: server error path (Jmax beyond the table) then a good request")
    c = W.config2(0)
    with Server(gpu, max_batch=8, streams=1) as srv:
        bad = dict(kind=0, pos1=c["pos1"], pos2=c["pos2"], L=70, grid_n=256, scale=0.3)
        import fastoverlap_amd as fo
        try:
            res = srv.align(bad)
            assert res["status"] != 0
        except fo.FastOverlapError:
            pass
        good = srv.align(dict(kind=0, pos1=c["pos1"], pos2=c["pos2"], groups=c["groups"],
                              L=c["L"], grid_n=c["grid_n"], scale=0.3))
        assert good["status"] == 0


def test_server_admits_by_hbm_footprint(gpu):
    """Admission by device footprint.  The budget holds about three of the
    burst's largest requests, far below the burst's total: the bytes the
    server's arenas hold never exceed it, requests wait for memory (several
    smaller batches instead of one), every result equals the request served
    alone, and a request larger than the whole budget fails with
    FO_ERR_NOMEM while the server keeps serving."""
    from fastoverlap_amd.batch import SphericalBatch
    pairs = [W.config2(p) for p in range(60)]
    sel = [c for c in pairs if c["grid_n"] == 64][:24]
    reqs = [dict(kind=0, pos1=c["pos1"], pos2=c["pos2"], groups=c["groups"], L=c["L"],
                 grid_n=64, scale=0.3) for c in sel]
    foot = []
    for c in sel:
        with SphericalBatch([len(c["pos1"])], [c["L"]], 0.3, 64, [c["groups"]]) as b:
            foot.append(b.footprint())
    budget = 3 * max(foot)
    assert budget < sum(foot) / 2
    rng = np.random.default_rng(1)
    X = rng.normal(size=(1000, 3)) * 3.0
    huge = dict(kind=0, pos1=X, pos2=X, L=30, grid_n=64, scale=0.3)
    say(f"case: server budget {budget} B ({len(reqs)} requests, footprints "
        f"{min(foot)}..{max(foot)} B, total {sum(foot)} B) + one 1000-atom request")
    with Server(gpu, max_batch=64, streams=3, max_wait_ms=2.0, hbm_budget=budget) as srv:
        ids = [srv.submit(r) for r in reqs[:12]]
        big = srv.submit(huge)
        ids += [srv.submit(r) for r in reqs[12:]]
        mixed = [srv.wait(i, timeout_ms=120000.0) for i in ids]
        hb = srv.wait(big, timeout_ms=120000.0)
        after = srv.align(reqs[0])
        st = srv.stats()
        alone = [srv.align(r) for r in reqs]
    assert hb["status"] == 3, hb["status"]          # FO_ERR_NOMEM
    assert after["status"] == 0
    assert st["peak_bytes"] <= budget, (st["peak_bytes"], budget)
    assert max(m["batch_size"] for m in mixed) < len(reqs)
    assert st["batches"] >= 3
    for k, (m, a) in enumerate(zip(mixed, alone)):
        assert m["status"] == 0 and a["status"] == 0, k
        assert m["dist"] == a["dist"], k
        np.testing.assert_array_equal(m["perm"], a["perm"])
        np.testing.assert_array_equal(m["ind"], a["ind"])
