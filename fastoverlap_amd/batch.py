"""Batched spherical / periodic alignment on the GPU (thin wrappers of the
``fo_sph_batch_*`` / ``fo_per_batch_*`` C ABI).

A batch is created once for a fixed set of shapes (atoms, Jmax, groups per
pair), keeps every buffer resident in device memory, and can be re-run on new
positions; ``run(graph=True)`` captures the launch sequence in a hipGraph on
the first call and replays it afterwards.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import (FastOverlapError, as_f64, as_i32, call, dptr, fptr, iptr)


# ---------------------------------------------------------------------------
# coefficient layouts
# ---------------------------------------------------------------------------

def kcount(L: int) -> int:
    """Number of (l, m1, m2) with l <= L."""
    return (L + 1) * (2 * L + 1) * (2 * L + 3) // 3


def compact_index(L: int):
    """(l, m1, m2) arrays in the compact l-major order of the C ABI."""
    ls, m1s, m2s = [], [], []
    for l in range(L + 1):
        m = np.arange(-l, l + 1)
        ls.append(np.full((2 * l + 1) ** 2, l))
        m1s.append(np.repeat(m, 2 * l + 1))
        m2s.append(np.tile(m, 2 * l + 1))
    return np.concatenate(ls), np.concatenate(m1s), np.concatenate(m2s)


def compact_to_dense(c, L: int, W: int | None = None):
    """compact complex (K(L),) -> dense (L+1, W, W) with m wrapped mod W
    (the reference's Ilmm layout, fastoverlap/sphericalAlignment.py:61-62)."""
    W = 2 * L + 1 if W is None else W
    l, m1, m2 = compact_index(L)
    out = np.zeros((L + 1, W, W), np.complex128)
    out[l, m1 % W, m2 % W] = c
    return out


def dense_to_compact(I):
    """dense (L+1, W, W) with m wrapped mod W -> compact complex (K(L),)."""
    L = I.shape[0] - 1
    W = I.shape[1]
    l, m1, m2 = compact_index(L)
    return np.ascontiguousarray(I[l, m1 % W, m2 % W], dtype=np.complex128)


def _groups_arrays(groups_per_pair):
    ngroups = as_i32([len(g) for g in groups_per_pair])
    sizes, atoms = [], []
    for g in groups_per_pair:
        for a in g:
            a = np.asarray(a, dtype=np.int64).ravel()
            sizes.append(len(a))
            atoms.append(a)
    sizes = as_i32(sizes if sizes else [0])
    atoms = as_i32(np.concatenate(atoms) if atoms else [0])
    return ngroups, sizes, atoms


def next_pow2(x: int) -> int:
    return 1 << max(0, (int(x) - 1).bit_length())


# ---------------------------------------------------------------------------
# spherical
# ---------------------------------------------------------------------------

def _harmonic_dense(c, L):
    """[groups][q(l,m), m >= 0][n] -> [groups][n][l][m % (2L+1)], negative m
    from c[l,-m] = (-1)^m conj(c[l,m])."""
    ng, _, nn = c.shape
    out = np.zeros((ng, nn, L + 1, 2 * L + 1), np.complex128)
    for l in range(L + 1):
        for m in range(l + 1):
            v = c[:, l * (l + 1) // 2 + m, :]
            out[:, :, l, m] = v
            if m:
                out[:, :, l, -m] = (-1) ** m * v.conj()
    return out


class SphericalBatch:
    """``npairs`` cluster pairs on one SO(3) grid of side ``grid_n``.

    natoms, L, scale: per pair (scalars broadcast).  ``groups``: per pair a
    list of index arrays (permutation groups); default one group of all
    atoms.  flags: ``_lib.SPH_*``.
    """

    def __init__(self, natoms, L, scale=None, grid_n=None, groups=None, flags=0, ctx=None,
                 harmonic=None):
        natoms = np.atleast_1d(np.asarray(natoms, dtype=np.int64))
        P = len(natoms)
        L = np.broadcast_to(np.asarray(L, dtype=np.int64), (P,))
        scale = np.broadcast_to(np.asarray(0.0 if scale is None else scale, float), (P,))
        scale = np.where(np.isnan(scale), 0.0, scale)
        if grid_n is None:
            grid_n = 2 * (int(L.max()) + 1)
        if groups is None:
            groups = [[np.arange(n)] for n in natoms]
        self.ctx = ctx or _lib.default_context()
        self.P, self.grid_n, self.flags = P, int(grid_n), flags
        self.natoms = natoms.astype(np.int64)
        self.L = L.astype(np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms)])
        self.norient = 1 if (flags & _lib.SPH_NO_INVERT) and not (flags & _lib.SPH_GENERAL) else 2
        self.koff = np.concatenate([[0], np.cumsum([kcount(int(x)) for x in self.L])])
        self._nat = as_i32(self.natoms)
        self._L = as_i32(self.L)
        self._scale = as_f64(scale)
        self._ng, self._gs, self._ga = _groups_arrays(groups)
        d = _lib.SphBatchDesc(P, self.grid_n, iptr(self._nat), iptr(self._L), dptr(self._scale),
                              iptr(self._ng), iptr(self._gs), iptr(self._ga), flags)
        h = C.c_void_p()
        # harmonic = (nmax, harmscale[, basis]): harmonic-oscillator basis
        # coefficients (basis 0: SphericalHarmonicAlign's, 1: the Fortran
        # module's HARMONICCOEFFS)
        self.harmonic = None if harmonic is None else (
            int(harmonic[0]), float(harmonic[1]), int(harmonic[2]) if len(harmonic) > 2 else 0)
        self._groups = [len(g) for g in groups]
        if self.harmonic is None:
            call("fo_sph_batch_create", self.ctx.handle, C.byref(d), C.byref(h))
        else:
            call("fo_sph_batch_create_harmonic", self.ctx.handle, C.byref(d),
                 self.harmonic[0], self.harmonic[1], self.harmonic[2], C.byref(h))
        self.handle = h

    # -- inputs ------------------------------------------------------------
    def _cat(self, pos):
        if isinstance(pos, (list, tuple)):
            pos = np.concatenate([np.asarray(p, float).reshape(-1, 3) for p in pos])
        pos = as_f64(pos).reshape(-1, 3)
        if len(pos) != self.offsets[-1]:
            raise ValueError(f"expected {self.offsets[-1]} atoms, got {len(pos)}")
        return pos

    def set_positions(self, pos1, pos2):
        a, b = self._cat(pos1), self._cat(pos2)
        call("fo_sph_batch_set_positions", self.handle, dptr(a), dptr(b))

    def set_positions_device(self, ptr1: int, ptr2: int):
        call("fo_sph_batch_set_positions_device", self.handle, C.c_void_p(ptr1), C.c_void_p(ptr2))

    def set_coeffs(self, A, B=None):
        A = np.ascontiguousarray(A, np.complex128).ravel()
        if len(A) != self.koff[-1]:
            raise ValueError("coefficient length mismatch")
        Bv = None if B is None else np.ascontiguousarray(B, np.complex128).ravel()
        call("fo_sph_batch_set_coeffs", self.handle, A.ctypes.data_as(_lib._pd),
             None if Bv is None else Bv.ctypes.data_as(_lib._pd))

    def harmonic_coeffs(self, pair=0):
        """Harmonic coefficients of one pair after the coefficient stage:
        (c1, c2), each [ngroups][nmax+1][L+1][2L+1] complex with m at
        index m % (2L+1) (the reference's calcHarmCoeff layout)."""
        if self.harmonic is None:
            raise ValueError("not a harmonic-basis batch")
        nmax = self.harmonic[0]
        L = int(self.L[pair])
        ng = self._groups[pair]
        nq = (L + 1) * (L + 2) // 2
        c1 = np.zeros((ng, nq, nmax + 1), np.complex128)
        c2 = np.zeros_like(c1)
        call("fo_sph_batch_harmonic_coeffs", self.handle, int(pair),
             c1.ctypes.data_as(_lib._pd), c2.ctypes.data_as(_lib._pd))
        return _harmonic_dense(c1, L), _harmonic_dense(c2, L)

    def coef_norm2(self):
        """Per pair: sum of |C^l_{m1 m2}|^2 over the full coefficient set."""
        out = np.zeros(self.P)
        call("fo_sph_batch_coef_norm2", self.handle, dptr(out))
        return out

    def set_start(self, R0):
        R0 = as_f64(R0).ravel()
        call("fo_sph_batch_set_start", self.handle, dptr(R0))

    def set_rotation(self, R):
        R = as_f64(R).ravel()
        call("fo_sph_batch_set_rotation", self.handle, dptr(R))

    # -- execution ---------------------------------------------------------
    def run(self, stages=_lib.STAGE_ALL, graph=False, sync=True):
        call("fo_sph_batch_run", self.handle, stages, 1 if graph else 0)
        if sync:
            self.sync()

    def sync(self):
        call("fo_sph_batch_sync", self.handle)

    def enable_timing(self, on=True):
        call("fo_sph_batch_timing", self.handle, 1 if on else 0, None)

    def last_timing(self):
        ms = np.zeros(8, np.float32)
        call("fo_sph_batch_timing", self.handle, -1, fptr(ms))
        return dict(zip(["prep", "ylm", "radial", "contract", "plane", "select", "rotopt",
                         "finish"], ms.tolist()))

    # -- outputs -------------------------------------------------------------
    def results(self):
        N = int(self.offsets[-1])
        dist = np.zeros(self.P)
        X1 = np.zeros((N, 3))
        X2 = np.zeros((N, 3))
        perm = np.zeros(N, np.int32)
        R = np.zeros((self.P, 3))
        rmat = np.zeros((self.P, 3, 3))
        inv = np.zeros(self.P, np.int32)
        call("fo_sph_batch_results", self.handle, dptr(dist), dptr(X1), dptr(X2), iptr(perm),
             dptr(R), dptr(rmat), iptr(inv))
        o = self.offsets
        return dict(dist=dist, R=R, rmat=rmat, inverted=inv.astype(bool),
                    X1=[X1[o[p]:o[p + 1]] for p in range(self.P)],
                    X2=[X2[o[p]:o[p + 1]] for p in range(self.P)],
                    perm=[perm[o[p]:o[p + 1]].astype(np.int64) for p in range(self.P)])

    def peaks(self):
        m = self.P * self.norient
        ind = np.zeros((m, 3), np.int32)
        fun, val, mar, dl = (np.zeros(m) for _ in range(4))
        nc, no = np.zeros(m, np.int32), np.zeros(m, np.int32)
        call("fo_sph_batch_peaks", self.handle, iptr(ind), dptr(fun), dptr(val), dptr(mar),
             dptr(dl), iptr(nc), iptr(no))
        sh = (self.P, self.norient)
        return dict(ind=ind.reshape(sh + (3,)), fun=fun.reshape(sh), value=val.reshape(sh),
                    margin=mar.reshape(sh), delta=dl.reshape(sh), ncand=nc.reshape(sh),
                    noverflow=no.reshape(sh))

    def precision(self):
        """Per (pair, orientation): the largest value of the reduced-precision
        grid, the FP64 value at the exact argmax and the error bound used."""
        m = self.P * self.norient
        v32, v64, dl = np.zeros(m), np.zeros(m), np.zeros(m)
        call("fo_sph_batch_precision", self.handle, dptr(v32), dptr(v64), dptr(dl))
        sh = (self.P, self.norient)
        return dict(val32=v32.reshape(sh), val64=v64.reshape(sh), delta=dl.reshape(sh))

    def tuning(self):
        """Launch shapes in use: (plane pairs per XCD chunk, a reserved entry
        without effect, radial workgroups per pair)."""
        v = np.zeros(3, np.int32)
        call("fo_sph_batch_tuning", self.handle, iptr(v), 0)
        return tuple(int(x) for x in v)

    def set_tuning(self, params):
        v = as_i32(params)
        call("fo_sph_batch_tuning", self.handle, iptr(v), 1)

    def start(self):
        """Interpolated grid maximum per (pair, orientation) as Euler angles
        (the refinement's start, fastoverlap/soft.py:127-130)."""
        R0 = np.zeros((self.P, self.norient, 3))
        call("fo_sph_batch_start", self.handle, dptr(R0))
        return R0

    def opt(self):
        m = self.P * self.norient
        R = np.zeros((m, 3))
        fun, gn = np.zeros(m), np.zeros(m)
        it = np.zeros(m, np.int32)
        ev = np.zeros(m, np.int32)
        call("fo_sph_batch_opt", self.handle, dptr(R), dptr(fun), dptr(gn), iptr(it), iptr(ev))
        sh = (self.P, self.norient)
        return dict(R=R.reshape(sh + (3,)), fun=fun.reshape(sh), gnorm=gn.reshape(sh),
                    iters=it.reshape(sh), evals=ev.reshape(sh))

    def scales(self):
        out = np.zeros(self.P)
        call("fo_sph_batch_scales", self.handle, dptr(out))
        return out

    def coeffs(self):
        out = np.zeros(int(self.koff[-1]), np.complex128)
        call("fo_sph_batch_coeffs", self.handle, out.ctypes.data_as(_lib._pd))
        return [out[self.koff[p]:self.koff[p + 1]] for p in range(self.P)]

    def grid(self, pair: int):
        n = self.grid_n
        g = np.zeros((self.norient, n, n, n), np.float32)
        call("fo_sph_batch_grid", self.handle, pair, fptr(g))
        return g

    def find_peaks(self, npeaks=10, width=2, residual=False):
        """Subtractive Gaussian peak search on every kept grid (needs
        SPH_KEEP_GRID and a run of STAGE_GRID).  Returns fit (P, norient,
        npeaks, 11): A, mu, the six s_ij, the centre as fractional grid
        indices; count (P, norient) of successful fits; optionally the
        residual grids (P, norient, n, n, n)."""
        P, no, n = self.P, self.norient, self.grid_n
        fit = np.zeros((P, no, npeaks, 11))
        cnt = np.zeros((P, no), np.int32)
        res = np.zeros((P, no, n, n, n), np.float32) if residual else None
        call("fo_sph_batch_find_peaks", self.handle, int(npeaks), int(width), dptr(fit),
             iptr(cnt), fptr(res))
        return fit, cnt, res

    def footprint(self) -> int:
        b = C.c_ulonglong()
        call("fo_sph_batch_footprint", self.handle, C.byref(b))
        return b.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fo_sph_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def align_spherical_batch(pos1, pos2, L, scale=None, grid_n=None, groups=None, invert=True,
                          ctx=None, diagnostics=False):
    """Align many cluster pairs in one batch.  pos1/pos2: lists of (N_p, 3)."""
    natoms = [len(np.asarray(p).reshape(-1, 3)) for p in pos1]
    Ls = np.broadcast_to(np.asarray(L), (len(natoms),))
    if grid_n is None:
        grid_n = next_pow2(2 * (int(Ls.max()) + 1))
    flags = 0 if invert else _lib.SPH_NO_INVERT
    with SphericalBatch(natoms, Ls, scale, grid_n, groups, flags, ctx) as b:
        b.set_positions(pos1, pos2)
        b.run()
        out = b.results()
        if diagnostics:
            out["peaks"] = b.peaks()
            out["opt"] = b.opt()
    return out


# ---------------------------------------------------------------------------
# periodic
# ---------------------------------------------------------------------------

def next_fast_len(target: int) -> int:
    """Smallest 2,3,5-smooth integer >= target (the FFT-size rule of
    fastoverlap/utils.py:278-313)."""
    if target <= 6:
        return target
    if not (target & (target - 1)):
        return target
    best = math.inf
    p5 = 1
    while p5 < target:
        p35 = p5
        while p35 < target:
            q = -(-target // p35)
            N = 2 ** ((q - 1).bit_length()) * p35
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


def periodic_defaults(natoms: int, boxvec, scale=None, n=None):
    """(scale, nk, F) as PeriodicAlign chooses them
    (fastoverlap/periodicAlignment.py:375-394)."""
    boxvec = np.asarray(boxvec, float)
    if scale is None:
        scale = (np.prod(boxvec) / natoms) ** (1.0 / 3) / 3.0
    nk = int(np.ceil(1.3 * natoms ** (1.0 / 3.0))) if n is None else int(n)
    F = next_fast_len(2 * (2 * nk + 1) + 1)
    return float(scale), nk, F


class PeriodicBatch:
    """``npairs`` periodic pairs sharing the k-cube half width ``nk`` and FFT
    side ``F``; per pair natoms, box (3,), scale."""

    def __init__(self, natoms, box, scale, nk, F, groups=None, niter=10, flags=0, ctx=None):
        natoms = np.atleast_1d(np.asarray(natoms, dtype=np.int64))
        P = len(natoms)
        box = np.broadcast_to(np.asarray(box, float).reshape(-1, 3) if np.ndim(box) > 1
                              else np.asarray(box, float), (P, 3))
        scale = np.broadcast_to(np.asarray(scale, float), (P,))
        if groups is None:
            groups = [[np.arange(n)] for n in natoms]
        self.ctx = ctx or _lib.default_context()
        self.P, self.nk, self.F, self.flags = P, int(nk), int(F), flags
        self.natoms = natoms
        self.offsets = np.concatenate([[0], np.cumsum(natoms)])
        self.ngroups_total = sum(len(g) for g in groups)
        self._nat = as_i32(natoms)
        self._box = as_f64(box)
        self._scale = as_f64(scale)
        self._ng, self._gs, self._ga = _groups_arrays(groups)
        d = _lib.PerBatchDesc(P, self.nk, self.F, iptr(self._nat), dptr(self._box),
                              dptr(self._scale), iptr(self._ng), iptr(self._gs), iptr(self._ga),
                              int(niter), flags)
        h = C.c_void_p()
        call("fo_per_batch_create", self.ctx.handle, C.byref(d), C.byref(h))
        self.handle = h

    def _cat(self, pos):
        if isinstance(pos, (list, tuple)):
            pos = np.concatenate([np.asarray(p, float).reshape(-1, 3) for p in pos])
        pos = as_f64(pos).reshape(-1, 3)
        if len(pos) != self.offsets[-1]:
            raise ValueError(f"expected {self.offsets[-1]} atoms, got {len(pos)}")
        return pos

    def set_positions(self, pos1, pos2):
        a, b = self._cat(pos1), self._cat(pos2)
        call("fo_per_batch_set_positions", self.handle, dptr(a), dptr(b))

    def set_positions_device(self, ptr1: int, ptr2: int):
        call("fo_per_batch_set_positions_device", self.handle, C.c_void_p(ptr1), C.c_void_p(ptr2))

    def set_coeffs(self, C1, C2):
        a = np.ascontiguousarray(C1, np.complex128)
        b = np.ascontiguousarray(C2, np.complex128)
        call("fo_per_batch_set_coeffs", self.handle, a.ctypes.data_as(_lib._pd),
             b.ctypes.data_as(_lib._pd))

    def set_disp(self, disp):
        d = as_f64(disp).ravel()
        call("fo_per_batch_set_disp", self.handle, dptr(d))

    def run(self, stages=_lib.PSTAGE_ALL, graph=False, sync=True):
        call("fo_per_batch_run", self.handle, stages, 1 if graph else 0)
        if sync:
            self.sync()

    def sync(self):
        call("fo_per_batch_sync", self.handle)

    def enable_timing(self, on=True):
        call("fo_per_batch_timing", self.handle, 1 if on else 0, None)

    def last_timing(self):
        ms = np.zeros(4, np.float32)
        call("fo_per_batch_timing", self.handle, -1, fptr(ms))
        return dict(zip(["sf", "plane", "select", "refine"], ms.tolist()))

    def results(self):
        N = int(self.offsets[-1])
        dist = np.zeros(self.P)
        X1 = np.zeros((N, 3))
        X2 = np.zeros((N, 3))
        perm = np.zeros(N, np.int32)
        disp = np.zeros((self.P, 3))
        call("fo_per_batch_results", self.handle, dptr(dist), dptr(X1), dptr(X2), iptr(perm),
             dptr(disp))
        o = self.offsets
        return dict(dist=dist, disp=disp,
                    X1=[X1[o[p]:o[p + 1]] for p in range(self.P)],
                    X2=[X2[o[p]:o[p + 1]] for p in range(self.P)],
                    perm=[perm[o[p]:o[p + 1]].astype(np.int64) for p in range(self.P)])

    def peaks(self):
        P = self.P
        ind = np.zeros((P, 3), np.int32)
        disp0 = np.zeros((P, 3))
        val, mar, dl = np.zeros(P), np.zeros(P), np.zeros(P)
        nc, no = np.zeros(P, np.int32), np.zeros(P, np.int32)
        call("fo_per_batch_peaks", self.handle, iptr(ind), dptr(disp0), dptr(val), dptr(mar),
             dptr(dl), iptr(nc), iptr(no))
        return dict(ind=ind, disp=disp0, value=val, margin=mar, delta=dl, ncand=nc, noverflow=no)

    def coeffs(self):
        NB = 2 * self.nk + 1
        C1 = np.zeros((self.ngroups_total, NB, NB, NB), np.complex128)
        C2 = np.zeros_like(C1)
        call("fo_per_batch_coeffs", self.handle, C1.ctypes.data_as(_lib._pd),
             C2.ctypes.data_as(_lib._pd))
        return C1, C2

    def product(self):
        NB = 2 * self.nk + 1
        Cp = np.zeros((self.P, NB, NB, NB), np.complex128)
        call("fo_per_batch_product", self.handle, Cp.ctypes.data_as(_lib._pd))
        return Cp

    def grid(self, pair: int):
        F = self.F
        g = np.zeros((F, F, F), np.float32)
        call("fo_per_batch_grid", self.handle, pair, fptr(g))
        return g

    def fft(self, pair: int):
        """Complex f = fftn(C, (F, F, F)) of one pair (needs PER_KEEP_FFT)."""
        F = self.F
        g = np.zeros((F, F, F, 2), np.float32)
        call("fo_per_batch_fft", self.handle, pair, fptr(g))
        return g[..., 0].astype(np.float64) + 1j * g[..., 1].astype(np.float64)

    def find_peaks(self, npeaks=10, width=2, residual=False):
        """Subtractive Gaussian peak search on every kept |f| grid (needs
        PER_KEEP_GRID and a run of PSTAGE_GRID): fit (P, npeaks, 11), count
        (P,), optionally the residual grids (P, F, F, F)."""
        P, F = self.P, self.F
        fit = np.zeros((P, npeaks, 11))
        cnt = np.zeros(P, np.int32)
        res = np.zeros((P, F, F, F), np.float32) if residual else None
        call("fo_per_batch_find_peaks", self.handle, int(npeaks), int(width), dptr(fit),
             iptr(cnt), fptr(res))
        return fit, cnt, res

    def set_coeffs_indexed(self, Cs, idx1, idx2):
        """Per-structure coefficients Cs (nstruct, ngroups, 2nk+1, 2nk+1,
        2nk+1) shared by the pairs: pair p uses structure idx1[p] as its first
        and idx2[p] as its second (PER_INPUT_COEFFS; every pair has the same
        groups)."""
        Cs = np.ascontiguousarray(Cs, np.complex128)
        i1 = as_i32(idx1)
        i2 = as_i32(idx2)
        call("fo_per_batch_set_coeffs_indexed", self.handle, Cs.ctypes.data_as(_lib._pd),
             int(Cs.shape[0]), iptr(i1), iptr(i2))

    def footprint(self) -> int:
        b = C.c_ulonglong()
        call("fo_per_batch_footprint", self.handle, C.byref(b))
        return b.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fo_per_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def align_periodic_batch(pos1, pos2, box, nk, F, scale, groups=None, niter=10, ctx=None,
                         diagnostics=False):
    natoms = [len(np.asarray(p).reshape(-1, 3)) for p in pos1]
    with PeriodicBatch(natoms, box, scale, nk, F, groups, niter, 0, ctx) as b:
        b.set_positions(pos1, pos2)
        b.run()
        out = b.results()
        if diagnostics:
            out["peaks"] = b.peaks()
    return out
