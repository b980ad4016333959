#!/usr/bin/env python3
"""Throughput of the MI355X structure-alignment stage on one GPU.

Workload (one step = the whole stage -- stage 1, stage 2 in two passes and
the final stage -- over one chunk): 16 query chains x 50,000 reference chains
from the seeded synthetic generator (lengths log-normal, median 250, sigma 0.5,
clipped to 30-1,000; 20% homologs as noisy rigid copies with indels), i.e.
800,000 query-reference pairs with the reference's default options.  Inputs
are resident in HBM when the timed region starts; results stay in HBM.

Prints one JSON line: pairs/s for the plain FP32 form (value) and the mixed
precision form, DP cell updates per second, the roofline fraction of the
dominant kernel class from HIP events around every launch of one profiled
step, and the reference's own CPU path (oracle/_ref) timed on a bounded
sample on this host, with the TM-score error of the GPU against it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from gtalign_alpha_amd import synth  # noqa: E402
from gtalign_alpha_amd.batch import GtaBatch  # noqa: E402
from gtalign_alpha_amd.capi import default_options, PREC_FP32, PREC_MIXED  # noqa: E402

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_form(qc, rc, prec, steps, warmup, profile):
    opts = default_options(precision=prec)
    with GtaBatch(0, opts, max_qrys=qc.nstrs, max_qry_poss=qc.nposs, max_refs=rc.nstrs,
                  max_ref_poss=rc.nposs) as b:
        b.TransferQueryPMDataToDevice(qc)
        b.TransferCPMDataToDevice(rc)
        for i in range(warmup):
            s, _ = b.process_resident()
            log("  warmup %d: %.3f s" % (i, s))
        t0 = time.perf_counter()
        dev, cells = 0.0, 0.0
        for i in range(steps):
            s, c = b.process_resident()
            dev += s
            cells += c
            log("  step %d: %.3f s" % (i, s))
        wall = time.perf_counter() - t0
        stats = None
        if profile:
            b.set_profiling(True)
            s, _ = b.process_resident()
            stats = b.kernel_stats()
            log("  profiled step: %.3f s" % s)
    return wall, dev, cells, stats


def cpu_baseline(qc, rc, gpu_tm_check=True):
    """The reference's own CPU path (oracle/_ref) on a bounded sample."""
    from oracle import oracle
    rng = np.random.default_rng(7)
    qsel = [0, qc.nstrs // 3, (2 * qc.nstrs) // 3, qc.nstrs - 1]
    nref = min(rc.nstrs, 2000)
    rsel = np.sort(rng.choice(rc.nstrs, nref, replace=False))
    q = qc.subset(qsel)
    r = rc.subset(rsel)
    if oracle.ref_available():
        threads = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        res = oracle.ref_process_block(q, r, nthreads=threads)
        dt = time.perf_counter() - t0
        kind = "reference"
    else:
        q = q.subset([0, 1])
        r = r.subset(list(range(min(400, r.nstrs))))
        threads = 1
        t0 = time.perf_counter()
        res = oracle.process_block(q, r)
        dt = time.perf_counter() - t0
        kind = "port"
    npairs = q.nstrs * r.nstrs
    out = {"value": npairs / dt, "unit": "pairs/s", "cores": threads, "kind": kind,
           "sample": "%d queries x %d references of the same workload (%d pairs), %.1f s" %
                     (q.nstrs, r.nstrs, npairs, dt)}
    if gpu_tm_check:
        with GtaBatch(0, default_options(precision=PREC_FP32), max_qrys=q.nstrs,
                      max_qry_poss=q.nposs, max_refs=r.nstrs, max_ref_poss=r.nposs) as b:
            b.TransferQueryPMDataToDevice(q)
            b.TransferCPMDataToDevice(r)
            g = b.ProcessBlock()
        aligned = res.alndata[..., 8] < 9999.0
        dtm = np.abs(g.alndata[..., 9:11] - res.alndata[..., 9:11])
        out["tm_err_max"] = float(dtm.max())
        out["aligned_pairs"] = int(aligned.sum())
        out["same_aligned_set"] = bool(np.array_equal(aligned, g.alndata[..., 8] < 9999.0))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the 50,000 reference chains (default: the full workload)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the stage's pairs shard across GPUs with no exchange "
              "(whole reference chunks per GPU), and runs on more than one GPU are not part of "
              "this build; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    qc, rc, _ = synth.config(1, scale=args.scale)
    npairs = qc.nstrs * rc.nstrs
    log("workload: %d queries x %d references = %d pairs (%d + %d residues), generated in %.1f s" %
        (qc.nstrs, rc.nstrs, npairs, qc.nposs, rc.nposs, time.perf_counter() - t0))

    log("plain FP32 form")
    wall, dev, cells, stats = time_form(qc, rc, PREC_FP32, args.steps, args.warmup, True)
    ms = 1e3 * wall / args.steps
    value = npairs * args.steps / wall
    mixed = None
    if not args.no_mixed:
        log("mixed-precision form")
        wall_m, dev_m, _, _ = time_form(qc, rc, PREC_MIXED, args.steps, args.warmup, False)
        mixed = npairs * args.steps / wall_m

    for st in stats:
        if st["launches"]:
            log("  %-16s launches %4d  %8.3f s  alg %.3f GB" % (st["name"], st["launches"], st["seconds"],
                                                            st["alg_bytes"] / 1e9))
    kern = {}
    dom = None
    for st in stats:
        if st["launches"] == 0:
            continue
        kern[st["name"]] = {"launches": st["launches"], "seconds": round(st["seconds"], 6),
                            "alg_GB": round(st["alg_bytes"] / 1e9, 3)}
        if dom is None or st["seconds"] > dom["seconds"]:
            dom = st
    achieved = dom["alg_bytes"] / dom["seconds"]
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                "kernel": dom["name"],
                "launch_ms": round(1e3 * dom["seconds"] / dom["launches"], 3),
                "alg_bytes_per_launch": dom["alg_bytes"] / dom["launches"]}

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(qc, rc)

    out = {
        "metric": "aligned pairs/sec + DP GCUPS on 1 MI355X, batched vs plain, TM err<=1e-4",
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "16 queries x %d references, lognormal(250,0.5) in [30,1000], "
                               "20%% homologs; default options" % rc.nstrs,
                   "pairs": npairs, "form": "plain FP32, one chunk per launch sequence"},
        "dp_gcups": round(cells / dev / 1e9, 2),
        "mixed_precision_pairs_per_s": round(mixed, 1) if mixed else None,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
