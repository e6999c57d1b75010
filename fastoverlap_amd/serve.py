"""Continuous-batching alignment server (wrapper of ``fo_server_*``).

Requests (spherical or periodic pairs) are copied at submission, queued by
shape class, admitted while their device footprint fits the budget, batched
and run on several HIP streams so that batches overlap on the device.
"""

from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import as_f64, as_i32, call


class Server:
    def __init__(self, ctx=None, max_batch=256, streams=2, max_wait_ms=2.0, hbm_budget=0):
        self.ctx = ctx or _lib.default_context()
        cfg = _lib.ServerConfig(int(max_batch), int(streams), float(max_wait_ms), int(hbm_budget))
        h = C.c_void_p()
        call("fo_server_create", self.ctx.handle, C.byref(cfg), C.byref(h))
        self.handle = h
        self._natoms = {}
        self._lock = threading.Lock()

    def submit(self, req: dict) -> int:
        """req: kind (0 spherical, 1 periodic), pos1, pos2, groups (list of
        index arrays, optional); spherical: L, grid_n (0 = next power of two),
        invert, scale (<= 0: automatic); periodic: box, nk, F, scale (<= 0 or
        absent: reference defaults)."""
        p1 = as_f64(req["pos1"]).reshape(-1, 3)
        p2 = as_f64(req["pos2"]).reshape(-1, 3)
        groups = req.get("groups")
        if groups is None:
            gs, ga, ng = as_i32([0]), as_i32([0]), 0
        else:
            gs = as_i32([len(g) for g in groups])
            ga = as_i32(np.concatenate([np.asarray(g).ravel() for g in groups]))
            ng = len(groups)
        r = _lib.Request()
        r.kind = int(req.get("kind", 0))
        r.natoms = len(p1)
        r.pos1 = _lib.dptr(p1)
        r.pos2 = _lib.dptr(p2)
        r.ngroups = ng
        r.group_sizes = _lib.iptr(gs)
        r.group_atoms = _lib.iptr(ga)
        r.L = int(req.get("L", 15))
        r.grid_n = int(req.get("grid_n", 0))
        r.invert = int(req.get("invert", 1))
        box = req.get("box", (1.0, 1.0, 1.0))
        for k in range(3):
            r.box[k] = float(np.asarray(box, float).ravel()[k])
        r.nk = int(req.get("nk", 0))
        r.F = int(req.get("F", 0))
        s = req.get("scale")
        r.scale = -1.0 if s is None else float(s)
        rid = C.c_longlong()
        call("fo_server_submit", self.handle, C.byref(r), C.byref(rid))
        with self._lock:
            self._natoms[rid.value] = len(p1)
        return rid.value

    def wait(self, rid: int, timeout_ms: float = -1.0) -> dict:
        with self._lock:
            n = self._natoms[rid]
        res = _lib.Result()
        X2 = np.zeros((n, 3))
        perm = np.zeros(n, np.int32)
        call("fo_server_wait", self.handle, rid, float(timeout_ms), C.byref(res), _lib.dptr(X2),
             _lib.iptr(perm))
        with self._lock:
            self._natoms.pop(rid, None)
        return dict(status=res.status, dist=res.dist, R=np.array(res.R[:]),
                    rmat=np.array(res.rmat[:]).reshape(3, 3), inverted=bool(res.inverted),
                    disp=np.array(res.disp[:]), ind=np.array(res.ind[:]),
                    latency_ms=res.latency_ms, queue_ms=res.queue_ms,
                    batch_size=res.batch_size, X2=X2, perm=perm.astype(np.int64))

    def align(self, req: dict) -> dict:
        return self.wait(self.submit(req))

    def stats(self) -> dict:
        c, b, p = C.c_longlong(), C.c_longlong(), C.c_ulonglong()
        call("fo_server_stats", self.handle, C.byref(c), C.byref(b), C.byref(p))
        return dict(completed=c.value, batches=b.value, peak_bytes=p.value)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fo_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
