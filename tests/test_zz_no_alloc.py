"""The allocation-free entry point fo_sph_align_noalloc against the
allocating batch path (fo_sph_batch_create + _set_positions + _run +
_results) on the same pairs: identical permutation and orientation,
distance within 1e-8 (the kernels and the layout are the same, so the
results are expected to be bit-identical), on two calls in a row, the second
with a smaller input.  Device buffers for the inputs and outputs are
allocated by the test itself (the caller's side), through the HIP runtime.
"""

import ctypes as C
import sys

import numpy as np
import pytest

from fastoverlap_amd import _lib, workloads as W
from fastoverlap_amd.batch import SphericalBatch, _groups_arrays

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes (caller-side buffers)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _desc(pairs, n, flags=0):
    nat = _lib.as_i32([len(c["pos1"]) for c in pairs])
    Ls = _lib.as_i32([c["L"] for c in pairs])
    sc = _lib.as_f64([c["scale"] for c in pairs])
    ng, gs, ga = _groups_arrays([c["groups"] for c in pairs])
    keep = (nat, Ls, sc, ng, gs, ga)
    d = _lib.SphBatchDesc(len(pairs), n, _lib.iptr(nat), _lib.iptr(Ls), _lib.dptr(sc),
                          _lib.iptr(ng), _lib.iptr(gs), _lib.iptr(ga), flags)
    return d, keep


def _noalloc(ctx, pairs, n):
    hip = _Hip()
    try:
        P = len(pairs)
        N = sum(len(c["pos1"]) for c in pairs)
        d1 = hip.put(np.concatenate([c["pos1"] for c in pairs]))
        d2 = hip.put(np.concatenate([c["pos2"] for c in pairs]))
        o_dist, o_X2 = hip.alloc(8 * P), hip.alloc(24 * N)
        o_perm, o_rmat, o_inv = hip.alloc(4 * N), hip.alloc(72 * P), hip.alloc(4 * P)
        d, keep = _desc(pairs, n)
        rc = _lib.load().fo_sph_align_noalloc(ctx.handle, C.byref(d), d1, d2, o_dist, o_X2,
                                              o_perm, o_rmat, o_inv)
        _lib.check(rc)
        return dict(dist=hip.get(o_dist, P, np.float64), X2=hip.get(o_X2, (N, 3), np.float64),
                    perm=hip.get(o_perm, N, np.int32), rmat=hip.get(o_rmat, (P, 3, 3), np.float64),
                    inverted=hip.get(o_inv, P, np.int32))
    finally:
        hip.free_all()


def _allocating(ctx, pairs, n):
    with SphericalBatch([len(c["pos1"]) for c in pairs], [c["L"] for c in pairs],
                        [c["scale"] for c in pairs], n, [c["groups"] for c in pairs], 0, ctx) as b:
        b.set_positions([c["pos1"] for c in pairs], [c["pos2"] for c in pairs])
        b.run()
        return b.results()


def _compare(tag, got, ref, pairs):
    off = 0
    for i, c in enumerate(pairs):
        N = len(c["pos1"])
        t = f"{tag} pair {i} N={N} L={c['L']}"
        np.testing.assert_array_equal(got["perm"][off:off + N], ref["perm"][i], err_msg=t)
        assert bool(got["inverted"][i]) == bool(ref["inverted"][i]), t
        assert abs(got["dist"][i] - ref["dist"][i]) <= 1e-8, (t, got["dist"][i], ref["dist"][i])
        np.testing.assert_allclose(got["X2"][off:off + N], ref["X2"][i], atol=1e-10, err_msg=t)
        np.testing.assert_allclose(got["rmat"][i], ref["rmat"][i], atol=1e-12, err_msg=t)
        off += N


def test_noalloc_matches_allocating_path_on_two_calls(gpu):
    ctx = gpu
    _lib.call("fo_sph_prepare", ctx.handle, 64)
    cand = [W.config2(p) for p in range(40)]
    big = [c for c in cand if c["grid_n"] == 64][:6]
    small = sorted((c for c in cand if c["grid_n"] == 64), key=lambda c: len(c["pos1"]))[:2]
    assert sum(len(c["pos1"]) for c in small) < sum(len(c["pos1"]) for c in big)
    for tag, pairs in (("call 1", big), ("call 2 (smaller)", small)):
        say(f"case: fo_sph_align_noalloc {tag}: {len(pairs)} pairs, "
            f"N={[len(c['pos1']) for c in pairs]}, grid 64")
        got = _noalloc(ctx, pairs, 64)
        ref = _allocating(ctx, pairs, 64)
        _compare(tag, got, ref, pairs)


def test_noalloc_refuses_what_it_cannot_hold(gpu):
    say("case: fo_sph_align_noalloc error paths (no tables; scratch too small)")
    # a fresh context has no resident tables: the entry point must not build them
    h = C.c_void_p()
    _lib.call("fo_context_create", 0, 0, C.byref(h))
    hip = _Hip()
    try:
        # both calls fail before any copy or launch; real buffers regardless
        buf = hip.alloc(24 * 1500)
        c = W.config0()
        d, keep = _desc([c], 32)
        rc = _lib.load().fo_sph_align_noalloc(h, C.byref(d), buf, buf,
                                              None, None, None, None, None)
        assert rc == 5, rc  # FO_ERR_STATE
        # one 1500-atom single-group pair at Jmax 15 needs ~72 MB of stored
        # radial values (four per atom pair) alone
        _lib.call("fo_sph_prepare", h, 32)
        rng = np.random.default_rng(0)
        X = rng.normal(size=(1500, 3)) * 4.0
        big = dict(pos1=X, pos2=X, L=15, scale=0.3, groups=[np.arange(1500)])
        d, keep = _desc([big], 32)
        rc = _lib.load().fo_sph_align_noalloc(h, C.byref(d), buf, buf,
                                              None, None, None, None, None)
        assert rc == 3, rc  # FO_ERR_NOMEM
        assert b"scratch" in _lib.load().fo_last_error()
    finally:
        hip.free_all()
        _lib.load().fo_context_destroy(h)
