"""Drop-in ``SOFT`` (``fastoverlap/soft.py:40-130``) on the GPU.

``iSOFT`` splits an arbitrary coefficient array f into two Hermitian parts,
f = H1 + i H2 with H^l_{-m1,-m2} = (-1)^{m1+m2} conj(H^l_{m1,m2}); the
inverse transform of a Hermitian set is real, so the HIP kernel computes the
two real grids in one pass and the result is iSOFT(H1) + i iSOFT(H2).
The grid is evaluated in FP32 (relative error well below 1e-5 of the grid
maximum); the Wigner table is built in FP64 on the device.
"""

from __future__ import annotations

import numpy as np
from numpy import pi

from . import _lib
from .batch import SphericalBatch, compact_to_dense, dense_to_compact, kcount


def hermitian_parts(flmm):
    """f -> (H1, H2), both Hermitian, with f = H1 + 1j*H2 (dense layout)."""
    bw = flmm.shape[0]
    W = flmm.shape[1]
    m = np.arange(W)
    m = np.where(m > W // 2, m - W, m)
    sign = (-1.0) ** (m[:, None] + m[None, :])
    neg = (-m) % W
    fh = sign[None] * np.conj(flmm[:, neg][:, :, neg])
    return 0.5 * (flmm + fh), (flmm - fh) / 2j


class SOFT:
    """SO(3) Fourier transform of bandwidth ``bw`` (grid side 2*bw)."""

    def __init__(self, bw):
        self.bw = bw
        self.bws = np.arange(0, bw * 2)
        self.n = 2 * bw
        self.Jmax = bw - 1
        self.weights = self.makeweights(bw)
        self.a = pi / bw * self.bws
        self.b = pi / 4 / bw * (2 * self.bws + 1)
        self.y = pi / bw * self.bws
        trip = [(l, m1, m2) for l in range(self.Jmax + 1)
                for m1 in range(-l, l + 1) for m2 in range(-l, l + 1)]
        self.Js, self.m1s, self.m2s = (np.array(v) for v in zip(*trip))
        self.indFactor = np.array([2 * pi / self.n, pi / self.n, 2 * pi / self.n])
        self._Ds = None

    @classmethod
    def makeweights(cls, bw):
        """Quadrature weights of the beta grid (``fastoverlap/soft.py:63-71``)."""
        j = np.arange(0, 2 * bw).astype(float)[:, None]
        k = np.arange(0, bw).astype(float)[None, :]
        fudge = pi / 4 / bw
        return (2 / (2 * k + 1) * np.sin((2 * j + 1) * (2 * k + 1) * fudge)
                * np.sin((2 * j + 1) * fudge) / bw).sum(1)

    @property
    def Ds(self):
        """Dense normalised Wigner table (bw, 2bw-1, 2bw-1, 2bw), FP64, built on
        the device by the three-term recurrence."""
        if self._Ds is None:
            bw = self.bw
            D = np.zeros((bw, 2 * bw - 1, 2 * bw - 1, 2 * bw))
            ctx = _lib.default_context()
            _lib.call("fo_sph_wigner_table", ctx.handle, bw, _lib.dptr(D))
            self._Ds = D
        return self._Ds

    def calcWignerMatrices(self):
        return self.Ds

    def iSOFT(self, flmm):
        """Inverse SOFT (``fastoverlap/soft.py:115-125``), complex (2bw)^3 grid."""
        bw = self.bw
        flmm = np.asarray(flmm, np.complex128)
        assert flmm.shape == (bw, bw * 2 - 1, bw * 2 - 1)
        H1, H2 = hermitian_parts(flmm)
        L = bw - 1
        with SphericalBatch([1], [L], [1.0], self.n, None,
                            _lib.SPH_GENERAL | _lib.SPH_KEEP_GRID) as b:
            b.set_coeffs(dense_to_compact(H1), dense_to_compact(H2))
            b.run(_lib.STAGE_GRID)
            g = b.grid(0).astype(np.float64)
        return g[0] + 1j * g[1]

    def SOFT(self, data):
        """Forward SOFT (``fastoverlap/soft.py:98-113``) on the device, FP64:
        2D DFT over alpha and gamma per beta plane, then the quadrature-
        weighted Wigner-d contraction over beta.  Returns (bw, 2bw-1, 2bw-1)."""
        bw = self.bw
        data = np.ascontiguousarray(np.asarray(data, np.complex128))
        assert all(n == self.n for n in data.shape)
        out = np.zeros(kcount(bw - 1), np.complex128)
        ctx = _lib.default_context()
        _lib.call("fo_soft_forward", ctx.handle, bw, data.ctypes.data_as(_lib._pd),
                  out.ctypes.data_as(_lib._pd))
        return compact_to_dense(out, bw - 1)

    def indtoEuler(self, ind):
        """Grid index -> Euler angles (``fastoverlap/soft.py:127-130``)."""
        R = self.indFactor * np.atleast_2d(np.asarray(ind, float))
        R[:, 1] += 0.5 * pi / self.n
        return R.squeeze()
