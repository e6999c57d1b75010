"""Write ``C2_pairs.npz``: the FP64 oracle's answers for the first 256 pairs of
the 4096-pair mixed cluster workload (N ~ U{13..147}, 1-2 species 70:30,
Jmax ~ U{15..25}, grid next_pow2(2(Jmax+1)), width 0.3).

The pairs themselves are regenerated from their seeds
(``fastoverlap_amd.workloads.config2``); a per-pair checksum of the positions
is stored so a test can tell a diverging generator from a wrong answer.
Stored per pair and orientation: the integer argmax, the objective at the
refined rotation, and the 8 largest real grid values with their flat
indices (so that a near-tie -- two grid points equal to FP64 rounding -- can
be judged by value instead of skipped); per pair: orientation choice,
permutation (concatenated) and distance.

    python tests/golden/make_c2_golden.py [--pairs 256] [--workers 8]
"""

import argparse
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fastoverlap_amd import workloads as W  # noqa: E402
from oracle import fo_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "C2_pairs.npz")
TOP = 8
_SOFT = {}


def checksum(c):
    return float(np.abs(c["pos1"]).sum() + 2.0 * np.abs(c["pos2"]).sum())


def one(p):
    c = W.config2(p)
    bw = c["grid_n"] // 2
    if bw not in _SOFT:
        _SOFT[bw] = O.SOFTOracle(bw)
    o = O.spherical_align(c["pos1"], c["pos2"], c["L"], c["scale"], perm=c["groups"], bw=bw,
                          fast=True, keep_grids=True, soft=_SOFT[bw])
    top_idx = np.zeros((2, TOP), np.int64)
    top_val = np.zeros((2, TOP))
    for k in range(2):
        g = o.grids[k].real.ravel()
        order = np.argsort(-g, kind="stable")[:TOP]
        top_idx[k] = order
        top_val[k] = g[order]
    return dict(p=p, L=c["L"], n=c["grid_n"], N=len(c["pos1"]), ind=np.array(o.ind),
                fun=np.array(o.fun), inverted=o.inverted, dist=o.dist,
                perm=np.asarray(o.perm, np.int64), top_idx=top_idx, top_val=top_val,
                chk=checksum(c))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--workers", type=int, default=8)
    a = ap.parse_args()
    with ProcessPoolExecutor(a.workers) as ex:
        rs = list(ex.map(one, range(a.pairs), chunksize=4))
    np.savez_compressed(
        OUT, pair=np.array([r["p"] for r in rs]), L=np.array([r["L"] for r in rs]),
        grid_n=np.array([r["n"] for r in rs]), natoms=np.array([r["N"] for r in rs]),
        ind=np.stack([r["ind"] for r in rs]), fun=np.stack([r["fun"] for r in rs]),
        inverted=np.array([r["inverted"] for r in rs]), dist=np.array([r["dist"] for r in rs]),
        perm=np.concatenate([r["perm"] for r in rs]), top_idx=np.stack([r["top_idx"] for r in rs]),
        top_val=np.stack([r["top_val"] for r in rs]), checksum=np.array([r["chk"] for r in rs]))
    Ls = sorted(set(r["L"] for r in rs))
    print(f"wrote {OUT}: {len(rs)} pairs, Jmax {Ls}, grids "
          f"{sorted(set(r['n'] for r in rs))}, inverted {sum(r['inverted'] for r in rs)}")


if __name__ == "__main__":
    main()
