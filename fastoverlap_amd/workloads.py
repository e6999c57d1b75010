"""Seeded synthetic inputs for the alignment workloads.

Every pair p of configuration c is drawn from
``numpy.random.default_rng(SeedSequence(2026, spawn_key=(c, p)))``.

Configurations (atom counts, radii, widths, degrees):
  C0 cluster      38 atoms, 1 species, radius 1.7, min distance 0.9,
                  sigma 0.3, Jmax 15, grid 32 (plain, FP64 path)
  C1 periodic     256 atoms (204 A + 52 B), cubic box 6.0, min distance
                  0.85, sigma 0.3, k-cube n = 8, FFT grid 64^3
  C2 cluster batch  N ~ U{13..147}, 1-2 species (70:30), Jmax ~ U{15..25},
                  grid = next power of two >= 2(Jmax+1), radius
                  1.2 * 1.717 * (N/38)^(1/3), 4096 pairs
  C3 large        1000 atoms (600/400) in radius 6.5, min distance 0.9,
                  sigma 0.2, Jmax 40, grid 128, 256 pairs
  C4 served mix   70% clusters (N 13..500, Jmax 15..30, radius as C2),
                  30% periodic (N 64..1024, 80:20, density 1.2, min distance
                  0.7, reference default n and sigma)

Uniform placement with rejection cannot reach C0 (38 atoms, R = 1.7,
d >= 0.9) or C1 (256 atoms, d >= 0.85): random sequential addition jams
first.  Those draws finish with a seeded soft-sphere relaxation (overlap and
wall penalties, steepest descent); if it does not reach the constraints the
radius (or box) grows by 1% steps and the value used is returned.
"""

from __future__ import annotations

import numpy as np

SEED = 2026


def rng_for(config: int, pair: int):
    return np.random.default_rng(np.random.SeedSequence(SEED, spawn_key=(config, pair)))


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------

def haar_rotation(rng):
    """Uniform random rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _uniform_ball(rng, n, R):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v * (R * rng.random(n) ** (1.0 / 3))[:, None]


def _min_image(d, box):
    return d - np.round(d / box) * box


def _rsa(rng, N, sample, dist2_fn, dmin, max_tries):
    """Random sequential addition: candidates are drawn in chunks of 256,
    pre-filtered against the accepted set in one vectorised step, then
    accepted in draw order."""
    X = np.empty((N, 3))
    n = 0
    tries = 0
    d2 = dmin * dmin
    while n < N and tries < max_tries:
        c = sample(256)
        tries += 256
        if n:
            c = c[dist2_fn(c[:, None, :] - X[None, :n, :]).min(1) >= d2]
        for x in c:
            if n and dist2_fn(X[:n] - x).min() < d2:
                continue
            X[n] = x
            n += 1
            if n == N:
                break
    return X[:n]


def _relax(X, dmin, R=None, box=None, iters=4000, step=0.15):
    """Steepest descent on sum relu(1.02 dmin - d)^2 (+ wall relu(|x|-R)^2),
    stopped as soon as the hard constraints (d >= dmin, |x| <= R) hold."""
    X = X.copy()
    target = 1.02 * dmin
    for _ in range(iters):
        d = X[:, None, :] - X[None, :, :]
        if box is not None:
            d = _min_image(d, box)
        r = np.linalg.norm(d, axis=2)
        np.fill_diagonal(r, np.inf)
        ov = np.clip(target - r, 0, None)
        if R is not None:
            rr = np.linalg.norm(X, axis=1)
            wall = np.clip(rr - R, 0, None)
        else:
            wall = 0.0
        if r.min() >= dmin and (R is None or not np.any(wall > 0)):
            break
        g = (ov / np.where(r > 0, r, 1))[:, :, None] * d
        force = g.sum(1)
        if R is not None:
            force -= (wall / np.where(rr > 0, rr, 1))[:, None] * X
        X += step * force
        if box is not None:
            X = X - np.floor(X / box) * box
    return X


def _ok_cluster(X, R, dmin):
    r = np.linalg.norm(X[:, None] - X[None], axis=2)
    np.fill_diagonal(r, np.inf)
    return r.min() >= dmin and np.linalg.norm(X, axis=1).max() <= R * (1 + 1e-9)


def _ok_periodic(X, box, dmin):
    d = _min_image(X[:, None] - X[None], box)
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    return r.min() >= dmin


def cluster(rng, N, R, dmin):
    """N points in a ball of radius R with pair distance >= dmin.
    Returns (X, R_used)."""
    X = _rsa(rng, N, lambda k: _uniform_ball(rng, k, R), lambda d: (d * d).sum(-1), dmin,
             400 * N)
    if len(X) == N:
        return X, R
    X = np.vstack([X, _uniform_ball(rng, N - len(X), R)])
    Ru = R
    for _ in range(80):
        X = _relax(X, dmin, R=Ru)
        if _ok_cluster(X, Ru, dmin):
            return X, Ru
        Ru *= 1.01
    raise RuntimeError("cluster generation failed")


def periodic_system(rng, N, box, dmin):
    """N points in a periodic box (3,) with min-image distance >= dmin."""
    box = np.asarray(box, float)
    def d2p(d):
        d = _min_image(d, box)
        return (d * d).sum(-1)
    X = _rsa(rng, N, lambda k: rng.random((k, 3)) * box, d2p, dmin, 100 * N)
    if len(X) == N:
        return X, box
    X = np.vstack([X, rng.random((N - len(X), 3)) * box])
    b = box.copy()
    for _ in range(80):
        X = _relax(X, dmin, box=b)
        if _ok_periodic(X, b, dmin):
            return X, b
        X *= 1.01
        b = b * 1.01
    raise RuntimeError("periodic generation failed")


def species_groups(N, nspecies, ratio=0.7):
    if nspecies == 1:
        return [np.arange(N)]
    na = int(round(ratio * N))
    return [np.arange(na), np.arange(na, N)]


def partner_cluster(rng, X, groups, noise=0.02):
    """Rotated (Haar), permuted within groups, N(0, noise) copy of X."""
    Rm = haar_rotation(rng)
    Y = X.dot(Rm.T)
    perm = np.arange(len(X))
    for g in groups:
        perm[g] = g[rng.permutation(len(g))]
    Y = Y[perm] + rng.normal(scale=noise, size=X.shape)
    return Y, perm, Rm


def partner_periodic(rng, X, box, groups, noise=0.02):
    """Translated (uniform in the box), permuted within groups, noisy copy."""
    t = rng.random(3) * box
    perm = np.arange(len(X))
    for g in groups:
        perm[g] = g[rng.permutation(len(g))]
    Y = X[perm] + t + rng.normal(scale=noise, size=X.shape)
    return Y, perm, t


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------

C0_RADIUS = 1.7
# Cluster radius scale of the batched / served mixes: R = 1.2 * 1.717 *
# (N/38)^(1/3).  1.717 is the C0 radius after relaxation; the factor 1.2
# makes plain rejection sampling succeed for every N in 13..500.
MIX_RADIUS = 1.2 * 1.717


def config0(pair=0):
    rng = rng_for(0, pair)
    X, Ru = cluster(rng, 38, C0_RADIUS, 0.9)
    X -= X.mean(0)
    g = [np.arange(38)]
    Y, perm, Rm = partner_cluster(rng, X, g)
    return dict(pos1=X, pos2=Y, groups=g, L=15, grid_n=32, scale=0.3, radius=Ru, true_perm=perm)


def config1(pair=0):
    rng = rng_for(1, pair)
    box = np.full(3, 6.0)
    X, b = periodic_system(rng, 256, box, 0.85)
    g = [np.arange(204), np.arange(204, 256)]
    Y, perm, t = partner_periodic(rng, X, b, g)
    return dict(pos1=X, pos2=Y, groups=g, box=b, scale=0.3, nk=8, F=64, shift=t, true_perm=perm)


def config2(pair, r0=MIX_RADIUS):
    rng = rng_for(2, pair)
    N = int(rng.integers(13, 148))
    ns = int(rng.integers(1, 3))
    L = int(rng.integers(15, 26))
    R = r0 * (N / 38.0) ** (1.0 / 3)
    X, _ = cluster(rng, N, R, 0.9)
    X -= X.mean(0)
    g = species_groups(N, ns)
    Y, perm, _ = partner_cluster(rng, X, g)
    grid_n = 1 << (2 * (L + 1) - 1).bit_length()
    return dict(pos1=X, pos2=Y, groups=g, L=L, grid_n=grid_n, scale=0.3, true_perm=perm)


def config3(pair):
    rng = rng_for(3, pair)
    X, _ = cluster(rng, 1000, 6.5, 0.9)
    X -= X.mean(0)
    g = species_groups(1000, 2, 0.6)
    Y, perm, _ = partner_cluster(rng, X, g)
    return dict(pos1=X, pos2=Y, groups=g, L=40, grid_n=128, scale=0.2, true_perm=perm)


def config4(pair, r0=MIX_RADIUS):
    """One request of the served mix (kind 0 spherical, 1 periodic)."""
    rng = rng_for(4, pair)
    if rng.random() < 0.7:
        N = int(rng.integers(13, 501))
        L = int(rng.integers(15, 31))
        R = r0 * (N / 38.0) ** (1.0 / 3)
        X, _ = cluster(rng, N, R, 0.9)
        X -= X.mean(0)
        g = species_groups(N, int(rng.integers(1, 3)))
        Y, perm, _ = partner_cluster(rng, X, g)
        grid_n = 1 << (2 * (L + 1) - 1).bit_length()
        return dict(kind=0, pos1=X, pos2=Y, groups=g, L=L, grid_n=grid_n, scale=0.3)
    N = int(rng.integers(64, 1025))
    box = np.full(3, (N / 1.2) ** (1.0 / 3))
    X, b = periodic_system(rng, N, box, 0.7)
    g = species_groups(N, 2, 0.8)
    Y, perm, t = partner_periodic(rng, X, b, g)
    return dict(kind=1, pos1=X, pos2=Y, groups=g, box=b)


def generate(config_fn, pairs, workers=None):
    """Draw many pairs of one configuration with a process pool."""
    pairs = list(pairs)
    if workers is None:
        import os
        workers = min(16, len(os.sched_getaffinity(0)))
    if workers <= 1 or len(pairs) < 8:
        return [config_fn(p) for p in pairs]
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(workers) as ex:
        return list(ex.map(config_fn, pairs, chunksize=max(1, len(pairs) // (4 * workers))))
