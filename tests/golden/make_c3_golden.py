"""Write the large-Jmax fixture ``C3_pairs.npz``: four 1000-atom two-species
(600:400) cluster pairs, Jmax 40, grid 128, Gaussian width 0.2, with the FP64
oracle's answer per pair (argmax per orientation, orientation, permutation,
distance and the 16 largest grid values of each orientation with their flat
indices).

The oracle takes ~40 s on one core per pair, too long to repeat inside the
GPU test, so its answers are stored here together with the inputs (the
positions themselves, so the fixture does not depend on the generator's
floating-point path on another host).

    python tests/golden/make_c3_golden.py [--pairs 4]
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "C3_pairs.npz")
TOP = 16


def one(p):
    c = W.config3(p)
    o = O.spherical_align(c["pos1"], c["pos2"], c["L"], c["scale"], perm=c["groups"],
                          bw=c["grid_n"] // 2, fast=True, keep_grids=True)
    top_idx = np.zeros((2, TOP), np.int64)
    top_val = np.zeros((2, TOP))
    gmax = np.zeros(2)
    for k in range(2):
        g = o.grids[k].real.ravel()
        order = np.argsort(-g, kind="stable")[:TOP]
        top_idx[k] = order
        top_val[k] = g[order]
        gmax[k] = np.abs(g).max()
    return dict(pos1=c["pos1"], pos2=c["pos2"], sizes=[len(g) for g in c["groups"]],
                ind=np.array(o.ind, np.int64), inverted=bool(o.inverted), dist=o.dist,
                perm=np.asarray(o.perm, np.int64), top_idx=top_idx, top_val=top_val, gmax=gmax)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=4)
    a = ap.parse_args()
    c0 = W.config3(0)
    with ProcessPoolExecutor(a.pairs) as ex:
        rs = list(ex.map(one, range(a.pairs)))
    np.savez(OUT, pos1=np.stack([r["pos1"] for r in rs]), pos2=np.stack([r["pos2"] for r in rs]),
             group_sizes=np.array([r["sizes"] for r in rs], np.int64),
             L=c0["L"], grid_n=c0["grid_n"], scale=c0["scale"],
             ind=np.stack([r["ind"] for r in rs]),
             inverted=np.array([r["inverted"] for r in rs]),
             dist=np.array([r["dist"] for r in rs]), perm=np.stack([r["perm"] for r in rs]),
             top_idx=np.stack([r["top_idx"] for r in rs]),
             top_val=np.stack([r["top_val"] for r in rs]), gmax=np.stack([r["gmax"] for r in rs]))
    for p, r in enumerate(rs):
        print(f"pair {p}: dist {r['dist']:.12f} ind {r['ind'].tolist()} inverted {r['inverted']}")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
