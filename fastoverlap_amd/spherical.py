"""Drop-in ``SphericalAlign`` and ``SphericalHarmonicAlign``
(``fastoverlap/sphericalAlignment.py:29-414``) running on the MI355X through
``libfastoverlap_hip``.

Same constructor and call signatures and return tuples as the reference.
Every numerical stage -- coefficients C^l (FP64), inverse SOFT (FP32 grid),
exact argmax (FP64 candidate re-evaluation), rotation refinement (FP64),
linear assignment and quaternion fit (FP64) -- runs in HIP kernels.

Differences from the reference, by design:
 * the inverted orientation uses C^l(X1, -X2) = (-1)^l C^l(X1, X2), an exact
   identity (Y_lm(-r) = (-1)^l Y_lm(r)), instead of recomputing;
 * the rotation refinement is a damped Newton iteration with the exact
   Hessian instead of scipy L-BFGS-B; both converge to the same local
   optimum, and only the permutation / RMSD downstream of it is contracted;
 * ``malign`` refines the inverted orientation's candidates from the
   inverted structure (the reference passes the non-inverted X2 there);
 * an atom at the centroid follows the i_0(0) = 1 limit instead of NaN;
 * ``SphericalHarmonicAlign.alignGroup`` aligns every pair exactly as
   ``align`` does (centred structures, each pair's own width, permutation
   groups); the reference builds the coefficients from the uncentred
   structures with the width of the previous call and one group
   (``sphericalAlignment.py:397-414``), which agrees when the structures are
   centred and the width is fixed.
"""

from __future__ import annotations

import math

import numpy as np
from numpy.linalg import norm

from . import _lib
from .batch import (SphericalBatch, compact_to_dense, dense_to_compact, kcount, next_pow2)
from .soft import SOFT, hermitian_parts
from .utils import EulerM, indtoEuler as _ind_to_euler


class _OptResult(dict):
    """Minimal scipy-OptimizeResult look-alike (``res.x``, ``res.fun``)."""

    __getattr__ = dict.__getitem__


def _groups_from_perm(perm, natoms):
    if perm is None:
        return [np.arange(natoms)]
    return [np.asarray(p, dtype=np.int64).ravel() for p in perm]


class BaseSphericalAlignment:
    calcScale = True

    def calcSO3Coeffs(self, pos1, pos2):  # pragma: no cover - abstract
        raise NotImplementedError

    def averageSeparation(self, pos):
        """Mean nearest-neighbour distance (``sphericalAlignment.py:35-41``),
        computed on the device by the same kernel the alignment uses."""
        pos = np.asarray(pos, float).reshape(-1, 3)
        with SphericalBatch([len(pos)], [0], [0.0], 2, None,
                            _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER) as b:
            b.set_positions(pos, pos)
            b.run(_lib.STAGE_PREP)
            s = b.scales()
        # the kernel returns (avg1 + avg2) / 6 with avg1 == avg2
        return float(s[0] * 3.0)

    def setJ(self, Jmax, bw=None):
        """``sphericalAlignment.py:44-55``.  ``bw`` (extension, default
        Jmax+1 as in the reference) sets the SOFT bandwidth: grid side 2*bw."""
        self.Jmax = Jmax
        self.J = np.arange(Jmax + 1)
        self.Jinds = np.array([j for j in range(Jmax + 1) for _ in range(-j, j + 1)])
        self.Minds = np.array([m for j in range(Jmax + 1) for m in range(-j, j + 1)])
        self.soft = SOFT(Jmax + 1 if bw is None else int(bw))
        trip = [(l, m1, m2) for l in range(Jmax + 1)
                for m1 in range(-l, l + 1) for m2 in range(-l, l + 1)]
        self.Js, self.m1s, self.m2s = (np.array(v) for v in zip(*trip))

    def sphHarm(self, theta, phi):
        """Y[l, m % (2J+1), ...] (``sphericalAlignment.py:57-65``) on the device."""
        theta = np.atleast_1d(np.asarray(theta, float))
        phi = np.atleast_1d(np.asarray(phi, float))
        assert theta.shape == phi.shape
        shape = theta.shape
        th = np.ascontiguousarray(theta.ravel())
        ph = np.ascontiguousarray(phi.ravel())
        L = self.Jmax
        Y = np.zeros((L + 1, 2 * L + 1, th.size), np.complex128)
        ctx = _lib.default_context()
        _lib.call("fo_sph_harm", ctx.handle, L, th.size, _lib.dptr(th), _lib.dptr(ph),
                  Y.ctypes.data_as(_lib._pd))
        return Y.reshape((L + 1, 2 * L + 1) + shape)

    def calcWignerMatrices(self, rot):
        """D(rot) and its gradient (``sphericalAlignment.py:67-96``)."""
        L = self.Jmax
        W = 2 * L + 1
        D = np.zeros((L + 1, W, W), np.complex128)
        G = np.zeros((3, L + 1, W, W), np.complex128)
        rot = _lib.as_f64(rot).ravel()
        ctx = _lib.default_context()
        _lib.call("fo_sph_wigner_D", ctx.handle, L, _lib.dptr(rot), D.ctypes.data_as(_lib._pd),
                  G.ctypes.data_as(_lib._pd))
        return D, G

    def getEnergyGradient(self, rot, Ilmm):
        """(-Re sum Ilmm D(rot), gradient) (``sphericalAlignment.py:98-100``)."""
        L = self.Jmax
        X = np.ascontiguousarray(Ilmm, np.complex128)
        rot = _lib.as_f64(rot).ravel()
        f = np.zeros(1)
        g = np.zeros(3)
        ctx = _lib.default_context()
        _lib.call("fo_sph_energy", ctx.handle, L, X.ctypes.data_as(_lib._pd), _lib.dptr(rot),
                  _lib.dptr(f), _lib.dptr(g))
        return float(f[0]), g

    def Hungarian(self, pos1, pos2, perm=None):
        """Per-group LAP on squared distances (``sphericalAlignment.py:102-108``).
        Returns (|X1 - X2[perm]|, perm)."""
        if perm is None:
            perm = self.perm if self.perm is not None else None
        X1 = _lib.as_f64(pos1).reshape(-1, 3)
        X2 = _lib.as_f64(pos2).reshape(-1, 3)
        return _lap(X1, X2, _groups_from_perm(perm, len(X1)), 0, None)

    def maxOverlap(self, R, Ilmm):
        """Local maximisation of Re sum Ilmm D(R) from R
        (``sphericalAlignment.py:110-113``); returns (x, res)."""
        L = self.Jmax
        Ic = dense_to_compact(np.conj(np.asarray(Ilmm)))  # kernel uses conj(I)
        with SphericalBatch([1], [L], [1.0], 2 * (L + 1), None, _lib.SPH_NO_INVERT) as b:
            b.set_coeffs(Ic)
            b.set_start(np.asarray(R, float).reshape(1, 3))
            b.run(_lib.STAGE_ROTOPT)
            o = b.opt()
        x = o["R"][0, 0]
        return x, _OptResult(x=x, fun=float(o["fun"][0, 0]), nit=int(o["iters"][0, 0]),
                             jac_norm=float(o["gnorm"][0, 0]))

    def rotate(self, X, R):
        return X.dot(EulerM(*R))

    def refine(self, X1, X2, R, permlist=None):
        """LAP + quaternion fit after rotating X2 by R (``:118-127``)."""
        if permlist is None:
            permlist = self.perm if self.perm is not None else None
        X1 = _lib.as_f64(X1).reshape(-1, 3)
        X2 = _lib.as_f64(X2).reshape(-1, 3)
        groups = [_groups_from_perm(permlist, len(X1))]
        with SphericalBatch([len(X1)], [0], [1.0], 2, groups,
                            _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER) as b:
            b.set_positions(X1, X2)
            b.set_rotation(np.asarray(R, float).reshape(1, 3))
            b.run(_lib.STAGE_PREP | _lib.STAGE_FINISH)
            r = b.results()
        return float(r["dist"][0]), X1, r["X2"][0]

    def COM_shift(self, pos1, pos2):
        """Centre both structures on their centroids (``:129-137``)."""
        X1 = np.array(pos1, float)
        X2 = np.array(pos2, float)
        X1 -= X1.mean(axis=0)[None, :]
        X2 -= X2.mean(axis=0)[None, :]
        return X1, X2

    # -- the hot path -------------------------------------------------------
    def _grid_n(self):
        return self.soft.n

    def _batch_kw(self):
        """Extra SphericalBatch arguments of the coefficient method."""
        return {}

    def align(self, pos1, pos2, perm=None, invert=True, calcCoeffs=None):
        """``sphericalAlignment.py:160-188``: returns (dist, X1, X2)."""
        return self._align(pos1, pos2, perm, invert, calcCoeffs)[:3]

    def _align(self, pos1, pos2, perm=None, invert=True, calcCoeffs=None):
        pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
        pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
        if perm is None:
            perm = self.perm
        groups = [_groups_from_perm(perm, len(pos1))]
        scale = 0.0 if self.calcScale else float(self.scale)
        L = self.Jmax
        if calcCoeffs is None:
            flags = 0 if invert else _lib.SPH_NO_INVERT
            b, reused = self._plain_batch(len(pos1), L, scale, groups, flags)
            b.set_positions(pos1, pos2)
            # a repeated shape replays the captured launch graph; a new one
            # runs eagerly (a capture costs more than one eager pass)
            b.run(_lib.STAGE_ALL, graph=reused)
            r = b.results()
            diag = dict(peaks=b.peaks(), opt=b.opt())
            if self.calcScale:
                self.scale = float(b.scales()[0])
        else:
            A = dense_to_compact(np.asarray(calcCoeffs(False)))
            flags = _lib.SPH_GENERAL if invert else _lib.SPH_NO_INVERT
            Bc = dense_to_compact(np.asarray(calcCoeffs(True))) if invert else None
            with SphericalBatch([len(pos1)], [L], [scale if scale > 0 else 1.0],
                                self._grid_n(), groups, flags) as b:
                b.set_positions(pos1, pos2)
                b.set_coeffs(A, Bc)
                b.run(_lib.STAGE_PREP | _lib.STAGE_GRID | _lib.STAGE_ROTOPT |
                      _lib.STAGE_FINISH)
                r = b.results()
                diag = dict(peaks=b.peaks(), opt=b.opt())
        out = (float(r["dist"][0]), r["X1"][0], r["X2"][0])
        self.last = dict(perm=r["perm"][0], R=r["R"][0], rmat=r["rmat"][0],
                         inverted=bool(r["inverted"][0]), **diag)
        return out

    def _plain_batch(self, n, L, scale, groups, flags):
        """The device batch of the last plain call and whether it is reused:
        it is kept while the shape repeats (atoms, degree, grid side, kernel
        width, flags, permutation groups), so repeated alignments of one
        system reuse its buffers and its captured launch graph.  Like the
        reference's instances (mutable scratch), an instance is not safe for
        concurrent calls."""
        kw = self._batch_kw()
        key = (n, L, self._grid_n(), scale, flags, tuple(sorted(kw.items())),
               tuple(tuple(int(i) for i in g) for g in groups[0]))
        cached = getattr(self, "_plain", None)
        if cached is not None and cached[0] == key:
            return cached[1], True
        if cached is not None:
            cached[1].close()
            self._plain = None
        b = SphericalBatch([n], [L], [scale], self._grid_n(), groups, flags, **kw)
        self._plain = (key, b)
        return b, False

    def findRotation(self, Ilmm):
        """iSOFT -> exact argmax -> refinement (``:190-194``): (R, fun)."""
        Ilmm = np.asarray(Ilmm, np.complex128)
        L = Ilmm.shape[0] - 1
        with SphericalBatch([1], [L], [1.0], max(self._grid_n(), 2 * (L + 1)), None,
                            _lib.SPH_NO_INVERT) as b:
            b.set_coeffs(dense_to_compact(Ilmm))
            b.run(_lib.STAGE_GRID | _lib.STAGE_ROTOPT)
            o = b.opt()
            self.last_peaks = b.peaks()
        return o["R"][0, 0], float(o["fun"][0, 0])

    def findRotations(self, Ilmm, nrot=10, width=2):
        """Euler angles of up to ``nrot`` Gaussian-fitted peaks of the real
        overlap grid of ``Ilmm`` (``sphericalAlignment.py:196-205``,
        ``utils.py:367-396``).  Returns (Rs, amplitude, mean, sigma, f) like
        the reference: the fitted peaks' heights, backgrounds and widths
        (2 s)^-1/2 of the six quadratic-form entries, and the residual grid."""
        Ilmm = np.asarray(Ilmm, np.complex128)
        L = Ilmm.shape[0] - 1
        n = max(self._grid_n(), 2 * (L + 1))
        H1, H2 = hermitian_parts(Ilmm)
        with SphericalBatch([1], [L], [1.0], n, None,
                            _lib.SPH_GENERAL | _lib.SPH_KEEP_GRID) as b:
            # the real part of iSOFT(f) is the grid of the Hermitian part H1
            b.set_coeffs(dense_to_compact(H1), dense_to_compact(H2))
            b.run(_lib.STAGE_GRID)
            fit, cnt, res = b.find_peaks(nrot, width, residual=True)
            pk = b.peaks()
            R0 = b.start()
        k = int(cnt[0, 0])
        if k == 0:
            # no fit succeeded: the interpolated maximum (utils.py:389-394)
            R = np.atleast_2d(R0[0, 0])
            return R, [float(pk["value"][0, 0])], [0.0], [np.nan], res[0, 0]
        f = fit[0, 0, :k]
        Rs = np.atleast_2d(_ind_to_euler(f[:, 8:11], n))
        with np.errstate(divide="ignore", invalid="ignore"):
            sig = list((2 * f[:, 2:8]) ** -0.5)
        return Rs, list(f[:, 0]), list(f[:, 1]), sig, res[0, 0]

    def malign(self, pos1, pos2, perm=None, invert=True, calcCoeffs=None, nrot=10):
        """Multi-peak alignment (``sphericalAlignment.py:206-235``): refine
        from every Gaussian-fitted peak of each orientation's grid (LAP +
        quaternion fit, no local maximisation) and keep the smallest
        distance.  Grids, peak fits and refinements run on the device.  The
        inverted orientation refines the inverted structure (the reference
        passes the non-inverted X2 to its inverted candidates)."""
        pos1 = np.asanyarray(pos1, float).reshape(-1, 3)
        pos2 = np.asanyarray(pos2, float).reshape(-1, 3)
        if perm is None:
            perm = self.perm
        groups = [_groups_from_perm(perm, len(pos1))]
        L = self.Jmax
        n = self._grid_n()
        if self.calcScale:
            self.scale = (self.averageSeparation(pos1) + self.averageSeparation(pos2)) / 6
        scale = float(self.scale)
        if calcCoeffs is None:
            flags = _lib.SPH_KEEP_GRID | (0 if invert else _lib.SPH_NO_INVERT)
            with SphericalBatch([len(pos1)], [L], [scale], n, groups, flags,
                                **self._batch_kw()) as b:
                b.set_positions(pos1, pos2)
                b.run(_lib.STAGE_PREP | _lib.STAGE_COEFF | _lib.STAGE_GRID)
                fit, cnt, _ = b.find_peaks(nrot, 2)
                R0 = b.start()
        else:
            A = dense_to_compact(np.asarray(calcCoeffs(False)))
            Bc = dense_to_compact(np.asarray(calcCoeffs(True))) if invert else None
            flags = _lib.SPH_KEEP_GRID | (_lib.SPH_GENERAL if invert else _lib.SPH_NO_INVERT)
            with SphericalBatch([len(pos1)], [L], [scale], n, groups, flags) as b:
                b.set_positions(pos1, pos2)
                b.set_coeffs(A, Bc)
                b.run(_lib.STAGE_GRID)
                fit, cnt, _ = b.find_peaks(nrot, 2)
                R0 = b.start()
        # candidate rotations per orientation (the interpolated maximum when
        # no fit succeeded, as findPeaks falls back to it)
        cands = []
        for o in range(fit.shape[1]):
            k = int(cnt[0, o])
            Rs = _ind_to_euler(fit[0, o, :k, 8:11], n) if k else R0[0, o][None, :]
            cands += [(o, np.asarray(R, float)) for R in np.atleast_2d(Rs)]
        P = len(cands)
        p2 = [(-pos2 if o == 1 else pos2) for o, _ in cands]
        with SphericalBatch([len(pos1)] * P, [0] * P, [scale] * P, 2, groups * P,
                            _lib.SPH_NO_INVERT) as fb:
            fb.set_positions([pos1] * P, p2)
            fb.set_rotation(np.array([R for _, R in cands]))
            fb.run(_lib.STAGE_PREP | _lib.STAGE_FINISH)
            r = fb.results()
        # first smallest within an orientation; the inverted orientation
        # replaces only on a strictly smaller distance
        best = 0
        for i in range(1, P):
            if r["dist"][i] < r["dist"][best]:
                best = i
        self.last_malign = dict(candidates=cands, dist=r["dist"].copy(), chosen=best,
                                inverted=cands[best][0] == 1, perm=r["perm"][best])
        return float(r["dist"][best]), r["X1"][best], r["X2"][best]

    def __call__(self, pos1, pos2, perm=None, invert=True, calcCoeffs=None, nrot=10, niter=None):
        """``sphericalAlignment.py:237-247``."""
        dist, X1, X2 = self.align(pos1, pos2, perm, invert, calcCoeffs)
        if (norm(X1 - X2, axis=1) > self.scale).sum() > len(X1) / 3:
            try:
                mdist, mX1, mX2 = self.malign(pos1, pos2, perm, invert, calcCoeffs, nrot)
                if mdist < dist:
                    return mdist, mX1, mX2
            except Exception:
                pass
        return dist, X1, X2


class SphericalAlign(BaseSphericalAlignment):
    """``SphericalAlign(scale=None, Jmax=15, perm=None)``
    (``sphericalAlignment.py:250-273``).  ``bw`` is an extension: the SOFT
    bandwidth (grid side 2*bw), default Jmax+1 as in the reference."""

    def __init__(self, scale=None, Jmax=15, perm=None, bw=None):
        if scale is not None:
            self.scale = scale
            self.calcScale = False
        else:
            self.calcScale = True
        self.setJ(Jmax, bw)
        self.perm = perm

    def calcSO3Coeffs(self, pos1, pos2):
        """C^l_{m1 m2} of two point sets, positions used as given
        (``sphericalAlignment.py:260-273``); dense (J+1, 2J+1, 2J+1)."""
        pos1 = np.atleast_2d(np.asarray(pos1, float))
        pos2 = np.atleast_2d(np.asarray(pos2, float))
        assert pos1.shape == pos2.shape
        L = self.Jmax
        with SphericalBatch([len(pos1)], [L], [float(self.scale)], 2 * (L + 1), None,
                            _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER) as b:
            b.set_positions(pos1, pos2)
            b.run(_lib.STAGE_PREP | _lib.STAGE_COEFF)
            c = b.coeffs()[0]
        return compact_to_dense(c, L)


def _lap(X1, X2, groups, metric, box):
    ng = len(groups)
    sizes = _lib.as_i32([len(g) for g in groups])
    atoms = _lib.as_i32(np.concatenate(groups) if ng else [0])
    perm = np.zeros(len(X1), np.int32)
    d = np.zeros(1)
    boxa = None if box is None else _lib.as_f64(box)
    ctx = _lib.default_context()
    _lib.call("fo_lap_align", ctx.handle, len(X1), _lib.dptr(X1), _lib.dptr(X2), ng,
              _lib.iptr(sizes), _lib.iptr(atoms), metric, _lib.dptr(boxa), _lib.iptr(perm),
              _lib.dptr(d))
    return float(d[0]), perm.astype(np.int64)


class SphericalHarmonicAlign(BaseSphericalAlignment):
    """``SphericalHarmonicAlign(scale=None, harmscale=1.0, nmax=15, Jmax=15,
    perm=None)`` (``sphericalAlignment.py:276-414``): the overlap coefficients
    come from each structure's projection on the harmonic-oscillator basis of
    radial degree <= nmax and length scale ``harmscale``,
    C^l = sum_n conj(c1) c2 (``:340-368``), built on the device by the
    batch's harmonic coefficient stage; the grid, argmax, rotation
    refinement, LAP and fit are the spherical pipeline's.  ``bw`` is an
    extension: the SOFT bandwidth (default Jmax+1)."""

    def __init__(self, scale=None, harmscale=1.0, nmax=15, Jmax=15, perm=None, bw=None):
        if scale is not None:
            self.scale = scale
            self.calcScale = False
        else:
            self.calcScale = True
        self.harmscale = harmscale
        self._bw = bw
        self.setCoeffs(nmax, Jmax, harmscale)
        self.perm = perm

    @classmethod
    def HarmCoeffs(cls, nmax, lmax, r0):
        """Basis polynomial coefficients [n][l][s] (``:320-326``): N_nl times
        the coefficient of r^s in r^l L_n^{l+1/2}(r^2) (``utils.py:414-427``),
        from the Laguerre closed form."""
        c = np.zeros((nmax + 1, lmax + 1, 2 * nmax + lmax + 1))
        for n in range(nmax + 1):
            for l in range(lmax + 1):
                a = l + 0.5
                norm = math.exp(0.5 * (math.log(2.0) + math.lgamma(n + 1) -
                                       (2 * l + 3) * math.log(r0) - math.lgamma(n + l + 1.5)))
                for k in range(n + 1):
                    v = 1.0
                    for i in range(k + 1, n + 1):
                        v *= a + i
                    v /= math.factorial(n - k) * math.factorial(k)
                    c[n, l, l + 2 * k] = (-v if k & 1 else v) * norm
        return c

    def setCoeffs(self, nmax=None, Jmax=None, harmscale=None):
        """``:328-338``."""
        if nmax is None:
            nmax = self.nmax
        else:
            self.nmax = nmax
        if Jmax is None:
            Jmax = self.Jmax
        else:
            self.setJ(Jmax, self._bw)
        if harmscale is None:
            harmscale = self.harmscale
        else:
            self.harmscale = harmscale
        self.coeffs = self.HarmCoeffs(nmax, Jmax, harmscale)

    def _batch_kw(self):
        return {"harmonic": (int(self.nmax), float(self.harmscale))}

    def calcHarmCoeff(self, pos):
        """c[n, l, m % (2J+1)] of Gaussians (width ``self.scale``) at ``pos``,
        positions used as given (``:340-356``), on the device."""
        pos = np.atleast_2d(np.asarray(pos, float))
        L = self.Jmax
        with SphericalBatch([len(pos)], [L], [float(self.scale)], 2 * (L + 1), None,
                            _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER, **self._batch_kw()) as b:
            b.set_positions(pos, pos)
            b.run(_lib.STAGE_PREP | _lib.STAGE_COEFF)
            c1, _ = b.harmonic_coeffs(0)
        return c1[0]

    def calcSO3Coeffs(self, pos1, pos2):
        """sum_n conj(c1) c2 of two point sets (``:358-361``), dense
        (J+1, 2J+1, 2J+1), on the device."""
        pos1 = np.atleast_2d(np.asarray(pos1, float))
        pos2 = np.atleast_2d(np.asarray(pos2, float))
        assert pos1.shape == pos2.shape
        L = self.Jmax
        with SphericalBatch([len(pos1)], [L], [float(self.scale)], 2 * (L + 1), None,
                            _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER, **self._batch_kw()) as b:
            b.set_positions(pos1, pos2)
            b.run(_lib.STAGE_PREP | _lib.STAGE_COEFF)
            c = b.coeffs()[0]
        return compact_to_dense(c, L)

    def alignGroup(self, coords, keepCoords=False):
        """Every ordered pair of ``coords`` aligned (``:397-414``), all n^2
        pairs as one device batch; returns dists (and the aligned pairs
        [2, n, n, N, 3] with keepCoords)."""
        coords = [np.asarray(c, float).reshape(-1, 3) for c in coords]
        n = len(coords)
        natoms = len(coords[0])
        if self.calcScale:
            # the reference computes the width of each pair from that pair
            seps = [self.averageSeparation(c) for c in coords]
        groups = [_groups_from_perm(self.perm, natoms)]
        P = n * n
        i1 = np.repeat(np.arange(n), n)
        i2 = np.tile(np.arange(n), n)
        scales = ([(seps[i] + seps[j]) / 6 for i, j in zip(i1, i2)] if self.calcScale
                  else [float(self.scale)] * P)
        with SphericalBatch([natoms] * P, [self.Jmax] * P, scales, self._grid_n(), groups * P,
                            0, **self._batch_kw()) as b:
            b.set_positions([coords[i] for i in i1], [coords[j] for j in i2])
            b.run(_lib.STAGE_ALL)
            r = b.results()
        dists = r["dist"].reshape(n, n)
        if keepCoords:
            aligned = np.empty((2, n, n, natoms, 3))
            for k in range(P):
                aligned[0, k // n, k % n] = r["X1"][k]
                aligned[1, k // n, k % n] = r["X2"][k]
            return dists, aligned
        return dists


class SphericalHarmonicAlignFortran(BaseSphericalAlignment):
    """``SphericalHarmonicAlignFortran(scale=0.3, Jmax=15, harmscale=1.0,
    nmax=20, perm=None, Natoms=None)`` (``sphericalAlignment.py:535-663``):
    the calls of the reference's f2py wrapper of ``ALIGNHARM`` and
    ``CALCOVERLAPMATRICES``, with the Fortran module's harmonic basis
    (``HARMONICCOEFFS``, ``f90/fastclusters.f90:492-687``: normalised
    oscillator functions of r / harmscale) built on the device.  As the
    library's Fortran cluster entry point, ``__call__`` refines the best
    overlap peak (``nrot`` is accepted for signature compatibility)."""

    def __init__(self, scale=0.3, Jmax=15, harmscale=1.0, nmax=20, perm=None, Natoms=None):
        self.scale = scale
        self.calcScale = False
        self.harmscale = harmscale
        self.nmax = nmax
        self.Natoms = Natoms
        self.setJ(Jmax)
        if perm is not None:
            self.setPerm(perm)
        elif Natoms is not None:
            self.setPerm([np.arange(Natoms)])
        else:
            self.perm = perm
        self.malign = self.__call__

    def setPerm(self, perm):
        self.Natoms = sum(map(len, perm))
        self.perm = perm

    def _batch_kw(self):
        return {"harmonic": (int(self.nmax), float(self.harmscale), 1)}

    def align(self, pos1, pos2, perm=None, invert=True, debug=False):
        return self(pos1, pos2, perm, invert, 1, debug)

    def __call__(self, pos1, pos2, perm=None, invert=True, nrot=10, debug=False):
        """(distance, pos1 centred, aligned pos2, rmatbest): rmatbest maps
        the (inverted, when chosen) centred pos2 onto the result."""
        if perm is not None:
            self.setPerm(perm)
        elif len(pos1) != self.Natoms:
            self.setPerm([np.arange(len(pos1))])
        dist, X1, X2 = BaseSphericalAlignment.align(self, pos1, pos2, self.perm, invert)
        # the batch's rmat maps the centred pos2 (rows) onto X2, inversion
        # included; RMATBEST acts on columns of the inverted input
        sg = -1.0 if self.last["inverted"] else 1.0
        rmat = np.asarray(self.last["rmat"], float).reshape(3, 3) / sg
        return dist, X1, X2, rmat.T.copy()

    def compareList(self, poslist, perm=None):
        """(avgoverlap, maxoverlap, navgoverlap, nmaxoverlap) of every pair
        of ``poslist`` (``sphericalAlignment.py:622-663``,
        ``CALCOVERLAPMATRICES``): sum |C|^2 and the maximum of the real
        overlap grid, all pairs i <= j as one device batch."""
        coords = np.array(poslist, float)
        nlist, Natoms, dim = coords.shape
        assert dim == 3
        if perm is None:
            if Natoms != self.Natoms:
                self.Natoms = Natoms
                self.setPerm([np.arange(self.Natoms)])
        else:
            self.setPerm(perm)
        coords -= coords.mean(1)[:, None, :]
        groups = [_groups_from_perm(self.perm, Natoms)]
        pairs = [(i, j) for i in range(nlist) for j in range(i, nlist)]
        P = len(pairs)
        L = self.Jmax
        with SphericalBatch([Natoms] * P, [L] * P, [float(self.scale)] * P, 2 * (L + 1),
                            groups * P, _lib.SPH_NO_INVERT | _lib.SPH_NO_CENTER,
                            **self._batch_kw()) as b:
            b.set_positions([coords[i] for i, _ in pairs], [coords[j] for _, j in pairs])
            b.run(_lib.STAGE_PREP | _lib.STAGE_COEFF | _lib.STAGE_GRID)
            norms = b.coef_norm2()
            maxs = b.peaks()["value"][:, 0]
        avg = np.zeros((nlist, nlist))
        mx = np.zeros((nlist, nlist))
        for k, (i, j) in enumerate(pairs):
            avg[i, j] = avg[j, i] = norms[k]
            mx[i, j] = mx[j, i] = maxs[k]
        da, dm = avg.diagonal(), mx.diagonal()
        return (avg, mx, avg / np.sqrt(da[:, None] * da[None, :]),
                mx / np.sqrt(dm[:, None] * dm[None, :]))
