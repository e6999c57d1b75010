"""Small host-side geometry helpers of the drop-in API.

These mirror helper functions of ``fastoverlap/utils.py`` that the drop-in
classes expose to callers (Euler matrices, angle conversion).  None of them
is a stage of the alignment pipeline, which runs entirely in HIP.
"""

from __future__ import annotations

import numpy as np
from numpy import cos, pi, sin


def EulerM(a, b, y):
    """Rotation matrix My . Mb . Ma (``fastoverlap/utils.py:447-460``)."""
    sa, ca = sin(a), cos(a)
    sb, cb = sin(b), cos(b)
    sy, cy = sin(y), cos(y)
    Ma = np.array(((ca, -sa, 0), (sa, ca, 0), (0, 0, 1)))
    Mb = np.array(((cb, 0, -sb), (0, 1, 0), (sb, 0, cb)))
    My = np.array(((cy, -sy, 0), (sy, cy, 0), (0, 0, 1)))
    return My.dot(Mb).dot(Ma)


def calcThetaPhiR(pos):
    """Polar angle, azimuth, radius (``fastoverlap/utils.py:439-445``)."""
    pos = np.atleast_2d(pos)
    X, Y, Z = pos.T
    R = np.linalg.norm(pos, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        theta = np.arccos(Z / R)
    return theta, np.arctan2(Y, X), R


def indtoEuler(ind, n):
    """Grid index -> Euler angles (``fastoverlap/utils.py:340-345``)."""
    ind = np.atleast_2d(np.asarray(ind, float))
    rot = np.array([2 * pi / n, pi / n, 2 * pi / n]) * ind
    rot[:, 1] += 0.5 * pi / n
    return rot.squeeze()
