"""Benchmark: aligned cluster pairs/s of the batched GPU path on the
4096-pair mixed workload, beside the plain per-call GPU drop-in, the FP64 CPU
oracle, the periodic path on C1 pairs, the large-degree C3 batch and a
served Poisson stream through the continuous-batching server under a 32 GiB
HBM admission budget (pairs/s, p50/p99 latency).

One "step" = one pass of the full alignment (coefficients, inverse SOFT,
exact argmax, rotation refinement, LAP, quaternion fit, both orientations)
over the 4096 pairs, inputs resident in HBM (one hipGraph replay per grid
class).  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--no-serve]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned pairs/s batched+served vs plain GPU (same argmax); p50/p99 latency"
FP32_PEAK = 157.3e12
HBM_PEAK = 8.0e12
FP64_PEAK = 78.6e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kcount(L):
    return (L + 1) * (2 * L + 1) * (2 * L + 3) // 3


def isoft_flops(L, n):
    """Algorithmic flops of the inverse SOFT of both orientations:
    beta contraction 2*4*n*K(L) + 2D inverse FFTs 2*10*n^3*log2(n)."""
    return 2 * 4 * n * kcount(L) + 2 * 10 * n ** 3 * math.log2(n)


def coeff_flops(groups, L):
    """GEMM-chain flops of C^l for one orientation: sum_s 4 N_s^2 (L+1)^2 + 8 N_s K(L)."""
    return sum(4 * len(g) ** 2 * (L + 1) ** 2 + 8 * len(g) * kcount(L) for g in groups)


def rotopt_eval_flops(L):
    """FP64 flops of one objective + gradient + Hessian evaluation of the
    rotation refinement at degree L (array of L): one Wigner-d recurrence
    step per stored coefficient with m1 >= 0 (sum_l (l+1)(2l+1) of them),
    32 flops each (d, d', d'' updates and six real accumulations)."""
    L = np.asarray(L)
    steps = (L + 1) * (L + 2) * (4 * L + 3) // 6
    return 32.0 * steps


def measured_traffic(kernel, npairs):
    """HBM bytes per step (one launch per grid class) of `kernel`, measured
    with rocprof FETCH_SIZE / WRITE_SIZE passes on this workload and recorded
    in profiles/roofline_traffic.json; None when not recorded for it."""
    path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    k = doc.get("kernels", {}).get(kernel)
    if k is None or doc.get("pairs") != npairs:
        return None
    return k["hbm_read_bytes_per_step"] + k["hbm_write_bytes_per_step"]


def make_batches(pairs, ctx, streams=1, flags=0):
    """One batch per grid class, each split into `streams` sub-batches with
    their own HIP streams so that their kernels overlap on the device."""
    from fastoverlap_amd.batch import SphericalBatch
    out = []
    for n in sorted(set(c["grid_n"] for c in pairs)):
        sel_all = [c for c in pairs if c["grid_n"] == n]
        k = max(1, min(streams, len(sel_all) // 256))
        for part in range(k):
            sel = sel_all[part::k]
            b = SphericalBatch([len(c["pos1"]) for c in sel], [c["L"] for c in sel],
                               [c["scale"] for c in sel], n, [c["groups"] for c in sel], flags,
                               ctx)
            out.append((n, sel, b))
    return out


def precision_report(batches):
    """Per (pair, orientation): |largest reduced-precision grid value - FP64
    value at the exact argmax| / |FP64 value| (the error stated per result),
    and the bound the selection used, relative the same way."""
    errs, bounds = [], []
    for _, _, b in batches:
        pr = b.precision()
        v64 = np.abs(pr["val64"])
        errs.append((np.abs(pr["val32"] - pr["val64"]) / v64).ravel())
        bounds.append((pr["delta"] / v64).ravel())
    e, d = np.concatenate(errs), np.concatenate(bounds)
    return {"max_rel_err_of_grid_max": float(e.max()),
            "median_rel_err_of_grid_max": float(np.median(e)),
            "max_rel_bound": float(d.max()), "median_rel_bound": float(np.median(d))}


def timed_steps(batches, steps, warmup):
    def step():
        for _, _, b in batches:
            b.run(graph=True, sync=False)
        for _, _, b in batches:
            b.sync()
    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    return (time.perf_counter() - t0) / steps * 1e3


def kernel_times(batches, steps):
    kt = {}
    for _, _, b in batches:
        b.enable_timing(True)
        for _ in range(steps):
            b.run(graph=False)
            for k, v in b.last_timing().items():
                kt[k] = kt.get(k, 0.0) + v / steps
        b.enable_timing(False)
    return {k: round(v, 4) for k, v in kt.items()}


def agreement(batches_a, batches_b):
    """Pairs whose argmax (both orientations), permutation and orientation
    are identical in two runs of the same pairs, and the largest distance
    difference."""
    same, tot, worst = 0, 0, 0.0
    for (_, _, a), (_, _, b) in zip(batches_a, batches_b):
        ra, rb, pa, pb = a.results(), b.results(), a.peaks(), b.peaks()
        for i in range(a.P):
            tot += 1
            same += int(np.array_equal(pa["ind"][i], pb["ind"][i])
                        and np.array_equal(ra["perm"][i], rb["perm"][i])
                        and ra["inverted"][i] == rb["inverted"][i])
            worst = max(worst, abs(ra["dist"][i] - rb["dist"][i]))
    return f"{same}/{tot}", worst


def beta_gemm_arm(pairs, ctx, steps, warmup, ref_batches):
    """The split-BF16x3 matrix-core arm of the beta contraction on the same
    pairs: step time, kernel times, per-result error and agreement with the
    FP32 arm (argmax, permutation and distance are exact in both: the FP64
    re-evaluation widens its window by the arm's error bound)."""
    from fastoverlap_amd import _lib
    batches = make_batches(pairs, ctx, 1, _lib.SPH_BETA_GEMM)
    for n, sel, b in batches:
        b.set_positions([c["pos1"] for c in sel], [c["pos2"] for c in sel])
    ms = timed_steps(batches, steps, warmup)
    kt = kernel_times(batches, steps)
    same, worst = agreement(ref_batches, batches)
    out = {"ms_per_step": ms, "pairs_per_s": len(pairs) / ms * 1e3, "kernel_ms": kt,
           "identical_argmax_perm_vs_fp32_arm": same, "max_abs_dist_diff_vs_fp32_arm": worst,
           "precision": precision_report(batches)}
    for _, _, b in batches:
        b.close()
    return out


def c3_supplement(c3pairs, ctx, steps):
    """The large-degree workload (1000-atom pairs, Jmax 40, grid 128, width
    0.2): both arms of the beta contraction, step time and per-result error,
    and their agreement."""
    from fastoverlap_amd import _lib
    out, runs = {}, []
    for arm, flags in (("fp32", 0), ("bf16x3", _lib.SPH_BETA_GEMM)):
        batches = make_batches(c3pairs, ctx, 1, flags)
        for n, sel, b in batches:
            b.set_positions([c["pos1"] for c in sel], [c["pos2"] for c in sel])
        ms = timed_steps(batches, steps, 1)
        out[arm] = {"ms_per_step": ms, "pairs_per_s": len(c3pairs) / ms * 1e3,
                    "kernel_ms": kernel_times(batches, 1),
                    "precision": precision_report(batches)}
        runs.append(batches)
    same, worst = agreement(runs[0], runs[1])
    out["identical_argmax_perm_between_arms"] = same
    out["max_abs_dist_diff_between_arms"] = worst
    for batches in runs:
        for _, _, b in batches:
            b.close()
    out["workload"] = (f"C3: {len(c3pairs)} pairs, N=1000 (600:400), Jmax 40, grid 128, "
                       f"sigma 0.2, radius 6.5")
    return out


def cpu_oracle_pair(c):
    """One pair through the FP64 oracle; returns its answer for the parity
    report (argmax per orientation, permutation, distance, orientation)."""
    from oracle import fo_oracle as O
    o = O.spherical_align(c["pos1"], c["pos2"], c["L"], c["scale"], perm=c["groups"],
                          bw=c["grid_n"] // 2, fast=False)
    return o.ind, np.asarray(o.perm), float(o.dist), bool(o.inverted)


def cpu_model():
    """The host CPU's model name (lscpu / /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _noop(_):
    return 0


def cpu_baseline(pairs, per_core):
    """The FP64 oracle (the reference's algorithm restated in numpy/scipy:
    naive einsum coefficients, per-(m1,m2) iSOFT loop, pocketfft, L-BFGS-B,
    exact LAP) on the host cores, one process per core, over the first
    per_core * cores pairs of the workload.  Returns the timing record and the
    oracle's answers (for the parity report against the batched GPU path)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    os.environ["OMP_NUM_THREADS"] = "1"
    cores = min(16, len(os.sched_getaffinity(0)))
    nsample = min(len(pairs), per_core * cores)
    sample = pairs[:nsample]
    # fresh interpreters: the workers must not inherit this process's GPU state
    with ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
        list(ex.map(_noop, range(cores)))          # start the workers first
        t0 = time.perf_counter()
        answers = list(ex.map(cpu_oracle_pair, sample, chunksize=1))
        wall = time.perf_counter() - t0
    rec = {"value": nsample / wall, "unit": "pairs/s", "cores": cores, "kind": "port",
           "cpu": cpu_model(),
           "sample": f"first {nsample} of the 4096 pairs, FP64 numpy/scipy oracle "
                     f"(naive einsum coefficients, per-(m1,m2) iSOFT, L-BFGS-B, exact LAP), "
                     f"{cores} processes x 1 thread, {wall:.1f} s wall"}
    return rec, answers


def oracle_parity(sample, answers, bres):
    """Batched GPU answers against the oracle's on the CPU-baseline sample:
    pairs with identical argmax (both orientations), permutation and
    orientation, and the largest distance difference."""
    same, worst = 0, 0.0
    for c, (ind, perm, dist, inv) in zip(sample, answers):
        d, p, gi, ginv = bres[id(c)]
        ok = (all(tuple(int(v) for v in gi[k]) == tuple(ind[k]) for k in range(len(ind)))
              and np.array_equal(p, perm) and bool(ginv) == inv)
        same += int(ok)
        worst = max(worst, abs(d - dist))
    return {"oracle_identical_argmax_perm": f"{same}/{len(answers)}",
            "max_abs_dist_err": worst}


SERVE_BUDGET = 32 << 30   # HBM admission budget of the served stream (bytes)


def _serve_stream(ctx, reqs, gaps):
    from fastoverlap_amd.serve import Server
    with Server(ctx, max_batch=256, streams=3, max_wait_ms=2.0,
                hbm_budget=SERVE_BUDGET) as srv:
        srv.align(reqs[0])  # warm the tables / code paths of both kinds
        srv.align(reqs[-1])
        ids = []
        t0 = time.perf_counter()
        tnext = t0
        for r, g in zip(reqs, gaps):
            tnext += g
            while time.perf_counter() < tnext:
                pass
            ids.append(srv.submit(r))
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        stats = srv.stats()
    lat = np.array([r["latency_ms"] for r in res])
    ok = sum(r["status"] == 0 for r in res)
    return {"pairs_per_s": ok / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_batch": float(np.mean([r["batch_size"] for r in res])),
            "failed": len(reqs) - ok, "batches": stats["batches"],
            "hbm_budget_bytes": SERVE_BUDGET, "peak_arena_bytes": stats["peak_bytes"]}


def served(ctx, reqs, seed=7):
    """The served mix (70% clusters N 13..500 Jmax 15..30, 30% periodic
    N 64..1024) through the continuous-batching server: first all requests
    at once (capacity), then a Poisson stream at 80% of that capacity."""
    n_requests = len(reqs)
    cap = _serve_stream(ctx, reqs, np.zeros(n_requests))
    rate = 0.8 * cap["pairs_per_s"]
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    out = _serve_stream(ctx, reqs, gaps)
    plain = plain_mix(reqs[:min(40, n_requests)])
    out.update({"requests": n_requests, "rate_hz": rate,
                "capacity_pairs_per_s": cap["pairs_per_s"],
                "capacity_p99_ms": cap["p99_ms"],
                "plain_sequential_pairs_per_s": plain,
                "speedup_vs_plain": out["pairs_per_s"] / plain})
    return out


def plain_mix(reqs):
    """Sequential plain drop-in calls (one pair per call) on served-mix
    requests: SphericalAlign / PeriodicAlign with the reference defaults."""
    from fastoverlap_amd.periodic import PeriodicAlign
    from fastoverlap_amd.spherical import SphericalAlign
    sph, per = {}, {}
    t0 = time.perf_counter()
    for r in reqs:
        if r["kind"] == 0:
            key = (r["L"], r["grid_n"])
            a = sph.get(key)
            if a is None:
                a = sph[key] = SphericalAlign(r["scale"], r["L"], bw=r["grid_n"] // 2)
            a.align(r["pos1"], r["pos2"], perm=r["groups"])
        else:
            a = PeriodicAlign(len(r["pos1"]), r["box"], r["groups"])
            a.align(r["pos1"], r["pos2"])
    return len(reqs) / (time.perf_counter() - t0)


def periodic_batched(ctx, pairs, steps, warmup, plain_sample):
    """Supplement: the periodic path on C1 pairs (256-atom 80:20 binary
    mixtures, box 6, k-cube n = 8, 64^3 translation grid), all pairs in one
    batch (one hipGraph replay per step), beside sequential plain
    PeriodicAlign calls on a sample with the same FFT side, and how many of
    those give the batch's displacement index and permutation."""
    from fastoverlap_amd.batch import PeriodicBatch
    from fastoverlap_amd.periodic import PeriodicAlign
    c0 = pairs[0]
    P = len(pairs)
    b = PeriodicBatch([len(c["pos1"]) for c in pairs], [c["box"] for c in pairs],
                      [c["scale"] for c in pairs], c0["nk"], c0["F"],
                      [c["groups"] for c in pairs], 10, 0, ctx)
    b.set_positions([c["pos1"] for c in pairs], [c["pos2"] for c in pairs])
    for _ in range(warmup):
        b.run(graph=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(graph=True, sync=False)
    b.sync()
    dt = (time.perf_counter() - t0) / steps
    b.enable_timing(True)
    b.run(graph=False)
    kt = b.last_timing()
    b.enable_timing(False)
    r, pk = b.results(), b.peaks()
    b.close()
    agree = 0
    t0 = time.perf_counter()
    for i, c in enumerate(pairs[:plain_sample]):
        pa = PeriodicAlign(len(c["pos1"]), c["box"], c["groups"], scale=c["scale"], n=c["nk"])
        pa.fshape = (c["F"],) * 3
        _, _, _, perm, _ = pa.align(c["pos1"], c["pos2"])
        agree += int(np.array_equal(pa.last["peaks"]["ind"][0], pk["ind"][i])
                     and np.array_equal(perm, r["perm"][i]))
    plain = plain_sample / (time.perf_counter() - t0)
    # algorithmic work: structure factors as separable phase products, 8 flops
    # per complex multiply-add, (2n+1)^3 k per atom, both structures (FP64);
    # the pruned transform: F planes of (2n+1)^3 + (2n+1)^2 F + (2n+1) F^2
    # complex multiply-adds (FP32)
    NB = 2 * c0["nk"] + 1
    F = c0["F"]
    f_sf = sum(2 * 8 * NB ** 3 * len(c["pos1"]) for c in pairs)
    f_plane = P * 8 * F * (NB ** 3 + NB * NB * F + NB * F * F)
    roof = {}
    for k, fl, peak, unit in (("sf", f_sf, FP64_PEAK, "FP64 vector ALU"),
                              ("plane", f_plane, FP32_PEAK / 2,
                               "FP32 vector ALU, single-issue (packed FP32 disabled)")):
        if kt.get(k, 0) > 0:
            ach = fl / (kt[k] * 1e-3)
            roof[k] = {"achieved_tflops": ach / 1e12, "peak_tflops": peak / 1e12,
                       "frac": ach / peak, "engine": unit}
    return {"workload": f"C1: {P} periodic pairs, N=256 (204:52), box 6.0, k-cube n=8, "
                        f"64^3 grid, sigma 0.3", "pairs": P, "ms_per_step": dt * 1e3,
            "roofline": roof,
            "pairs_per_s": P / dt, "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
            "plain_pairs_per_s": plain, "speedup_vs_plain": (P / dt) / plain,
            "plain_identical_ind_perm": f"{agree}/{plain_sample}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-per-core", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--streams", type=int, default=1,
                    help="sub-batches (HIP streams) per grid class")
    ap.add_argument("--gen-workers", type=int, default=None,
                    help="processes drawing the synthetic pairs (default: up to 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true", help="(default) run the served mix")
    ap.add_argument("--no-serve", action="store_true", help="skip the served mix")
    ap.add_argument("--periodic-pairs", type=int, default=128,
                    help="C1 pairs of the periodic supplement (0: skip it)")
    ap.add_argument("--no-periodic", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the default launch shapes instead of the autotuned ones")
    ap.add_argument("--c3-pairs", type=int, default=256,
                    help="pairs of the large-degree supplement (0: skip it)")
    ap.add_argument("--no-bf16", action="store_true",
                    help="skip the split-BF16x3 arm of the beta contraction")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(pairs are independent; see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    from fastoverlap_amd import _lib, workloads as W
    from fastoverlap_amd.spherical import SphericalAlign

    t = time.perf_counter()
    pairs = W.generate(W.config2, range(args.pairs), workers=args.gen_workers)
    # every workload is drawn (with forked workers) before the GPU is touched
    run_serve = not args.no_serve and args.serve_requests > 0
    serve_reqs = ([] if not run_serve else
                  W.generate(W.config4, range(args.serve_requests), workers=args.gen_workers))
    nper = 0 if args.no_periodic else args.periodic_pairs
    per_pairs = W.generate(W.config1, range(nper), workers=args.gen_workers) if nper else []
    c3_pairs = (W.generate(W.config3, range(args.c3_pairs), workers=args.gen_workers)
                if args.c3_pairs else [])
    log(f"generated {len(pairs)} pairs, {len(per_pairs)} periodic pairs and "
        f"{len(serve_reqs)} served requests in "
        f"{time.perf_counter() - t:.1f}s")
    ctx = _lib.default_context()
    batches = make_batches(pairs, ctx, args.streams,
                           0 if args.no_autotune else _lib.SPH_AUTOTUNE)
    for n, sel, b in batches:
        b.set_positions([c["pos1"] for c in sel], [c["pos2"] for c in sel])
        log(f"grid {n}: {len(sel)} pairs, {b.footprint() / 2**20:.0f} MiB on device")

    def step(graph=True):
        for _, _, b in batches:
            b.run(graph=graph, sync=False)
        for _, _, b in batches:
            b.sync()

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = len(pairs) * args.steps / elapsed
    log(f"batched: {ms_per_step:.2f} ms/step, {value:.0f} pairs/s")

    # the autotuner's choice per grid class, and the default shapes timed on
    # the same batches for comparison
    tuning = {"autotune": not args.no_autotune,
              "shapes": {str(n): b.tuning() for n, _, b in batches}}
    if not args.no_autotune:
        chosen = [b.tuning() for _, _, b in batches]
        for _, _, b in batches:
            b.set_tuning((16, 15, 0))
        step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        tuning["ms_per_step_default_shapes"] = (time.perf_counter() - t0) / args.steps * 1e3
        for (_, _, b), sh in zip(batches, chosen):
            b.set_tuning(sh)
        step()
        log(f"autotune: {tuning}")

    # per-kernel times: HIP events recorded on each batch's stream between the
    # stages, averaged over `steps` eager passes of the same work
    kt = {}
    for n, sel, b in batches:
        b.enable_timing(True)
        for _ in range(args.steps):
            b.run(graph=False)
            for k, v in b.last_timing().items():
                kt[k] = kt.get(k, 0.0) + v / args.steps
        b.enable_timing(False)
    log("kernel ms per step: " + ", ".join(f"{k}={v:.3f}" for k, v in kt.items()))

    # batched results kept for the plain-path comparison
    bres = {}
    its, evs, Ls = [], [], []
    for n, sel, b in batches:
        o = b.opt()
        its.append(o["iters"].ravel())
        evs.append(o["evals"])
        Ls.append(np.array([c["L"] for c in sel]))
    its = np.concatenate(its)
    log(f"rotation refinement: Newton iterations mean {its.mean():.2f}, max {its.max()}; "
        f"evaluations mean {np.concatenate([e.ravel() for e in evs]).mean():.2f}")
    for n, sel, b in batches:
        r = b.results()
        pk = b.peaks()
        for i, c in enumerate(sel):
            bres[id(c)] = (r["dist"][i], r["perm"][i], pk["ind"][i], r["inverted"][i])

    # plain per-call drop-in on a sample
    sample = pairs[:args.plain_sample]
    aligners = {}
    agree = 0
    t0 = time.perf_counter()
    for c in sample:
        key = (c["L"], c["grid_n"])
        sa = aligners.get(key)
        if sa is None:
            sa = aligners[key] = SphericalAlign(c["scale"], c["L"], bw=c["grid_n"] // 2)
        sa.align(c["pos1"], c["pos2"], perm=c["groups"])
        d, p, ind, _ = bres[id(c)]
        agree += int(np.array_equal(sa.last["peaks"]["ind"][0], ind)
                     and np.array_equal(sa.last["perm"], p))
    plain_elapsed = time.perf_counter() - t0
    plain = len(sample) / plain_elapsed
    log(f"plain drop-in: {plain:.1f} pairs/s; identical argmax+perm {agree}/{len(sample)}")

    # roofline of the dominant kernel (algorithmic flops / measured time)
    work = {
        "plane": ("k_isoft_fft", sum(isoft_flops(c["L"], c["grid_n"]) for c in pairs), FP32_PEAK),
        "contract": ("k_sph_contract_low+k_sph_contract_mfma",
                     sum(coeff_flops(c["groups"], c["L"]) for c in pairs), FP64_PEAK),
        "rotopt": ("k_sph_rotopt", sum(float((e * rotopt_eval_flops(L)[:, None]).sum())
                                       for e, L in zip(evs, Ls)), FP64_PEAK),
    }
    per_kernel = {}
    for k, (name, fl, peak) in work.items():
        if kt.get(k, 0) > 0:
            ach = fl / (kt[k] * 1e-3)
            per_kernel[name] = {"ms": round(kt[k], 4), "achieved_tflops": ach / 1e12,
                                "peak_tflops": peak / 1e12, "frac": ach / peak}
    # the dominant kernel: the stage with the largest time, except that the
    # contraction stage is three kernels (degrees 0..7, 8..15 and >= 16),
    # each well below the plane kernel, so it is not one kernel to roofline
    dom = max((k for k in kt if k != "contract"), key=kt.get)
    # compute-bound kernels: the engine whose published peak bounds each one
    # ("valu": vector ALU, "mfma": matrix cores)
    bound = {"plane": "valu", "contract": "mfma", "rotopt": "valu"}
    compute = {"plane": "FP32 vector ALU; the published 157.3 TFLOP/s counts packed "
                        "(dual) FP32 issue, which this build keeps disabled, so 78.6 "
                        "TFLOP/s is its reachable single-issue ceiling",
               "contract": "FP64 matrix cores (v_mfma_f64_16x16x4f64; degrees <= 15 with the "
                           "radial recurrence fused in)",
               "rotopt": "FP64 vector ALU"}
    if dom in work:
        name, fl, peak = work[dom]
        ach = fl / (kt[dom] * 1e-3)
        roof = {"kernel": name, "bound": bound[dom], "compute": compute[dom],
                "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": ach / peak, "traffic": measured_traffic(name, len(pairs))}
        if dom == "plane":
            roof["frac_of_single_issue_peak"] = ach / (FP32_PEAK / 2)
    else:
        roof = {"kernel": dom, "bound": None, "achieved": None, "peak": None,
                "unit": "TFLOP/s", "frac": None, "traffic": None}
    roof["kernel_ms"] = {k: round(v, 4) for k, v in kt.items()}
    roof["per_kernel"] = per_kernel

    out = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "C2: 4096 cluster pairs, N~U{13..147}, 1-2 species 70:30, "
                               "Jmax~U{15..25}, grid next_pow2(2(Jmax+1)), sigma 0.3, "
                               "both orientations, inputs resident in HBM",
                   "pairs": len(pairs), "grids": {str(n): sum(len(s) for m, s, _ in batches if m == n)
                             for n in sorted(set(m for m, _, _ in batches))},
                   "streams_per_class": args.streams},
        "plain_gpu_pairs_per_s": plain,
        "speedup_vs_plain": value / plain,
        "plain_identical_argmax_perm": f"{agree}/{len(sample)}",
        "roofline": roof,
        "tuning": tuning,
        "precision": {"fp32_arm": precision_report(batches)},
    }
    if not args.no_bf16:
        out["precision"]["bf16x3_arm"] = beta_gemm_arm(pairs, ctx, args.steps, args.warmup,
                                                       batches)
        log(f"bf16x3 arm: {out['precision']['bf16x3_arm']}")
    if per_pairs:
        out["periodic"] = periodic_batched(ctx, per_pairs, args.steps, args.warmup,
                                           min(4, len(per_pairs)))
        log(f"periodic: {out['periodic']}")
    if c3_pairs:
        out["c3"] = c3_supplement(c3_pairs, ctx, 2)
        log(f"c3: {out['c3']}")
    if run_serve:
        out["serve"] = served(ctx, serve_reqs)
        log(f"served: {out['serve']}")
    if not args.no_cpu:
        rec, answers = cpu_baseline(pairs, args.cpu_per_core)
        out["cpu_baseline"] = rec
        out["oracle_parity"] = oracle_parity(pairs[:len(answers)], answers, bres)
        log(f"cpu baseline: {rec}; parity {out['oracle_parity']}")
    for _, _, b in batches:
        b.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
