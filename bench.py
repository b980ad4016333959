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

# After the stage-1 prescreen the reference keeps a reference for every query
# of a block when one query of the block keeps it; its scheduler forms blocks
# of --dev-queries-per-chunk queries (default 2, CLOptions.cpp:88).  The
# bench's calls use the same blocks, so they compute what the reference
# computes for the workload with default options.
QGROUP = 2

HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12      # MI355X FP32 vector (and dense matrix) peak
# the VALU issue ceiling that peak implies: one 64-lane FP32 multiply-add
# (2 x 64 operations) per SIMD every 2 cycles, 1,024 SIMDs at 2.4 GHz
VALU_ISSUE_PEAK = FP32_PEAK / (2 * 64)
# rocprofv3 SQ counter summary (VALU issue) of the same workload's kernels
VALU_PROFILE = "profiles/refine_dp_16x50000_valu_issue.md"
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the 16 x 50,000 workload's kernels
TRAFFIC_PROFILE = "profiles/refine_dp_16x50000_hbm_traffic.md"
# kernel class -> the kernel (template instance) rocprofv3 lists for it in the plain FP32 form
KERNEL_NAMES = {
    "findfrag": "gta::k_findfrag<0>",
    "refine_gapless": "gta::k_refine_gapless<0>",
    "dp": "gta::k_dp<false> (29 of 30 launches; 1 is k_dp<true>)",
    "refine_dp": "gta::k_refine_dp<0, 0, 1> (RM_DPREFINE, 1-wave workgroups on the bench's chunks)",
    "refine_final": "gta::k_refine_dp<0, 1..4> (final and production phases)",
}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` on the full 16 x 50,000 workload: the
    committed rocprofv3 counter summary's figure (FETCH_SIZE and WRITE_SIZE
    per dispatch, each scaled by what the counter reads for a copy of known
    size); None when the summary has no line for the kernel."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and " ".join(parts[1:-1]) == kernel:
                return int(parts[-1])
    return None


def profiled_valu(kernel):
    """VALU-issue fraction of `kernel` on the full workload from the committed
    SQ counter summary: SQ_INSTS_VALU per dispatch (wave-instructions, exactly
    what the counter reads for the calibration kernel of known instruction
    count) per second of the dispatch, against VALU_ISSUE_PEAK; None when
    absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), VALU_PROFILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "valu_issue_frac" and " ".join(parts[1:-1]) == kernel:
                return float(parts[-1])
    return None


def valu_roofline(dom, kname, full):
    """Roofline of the dominant kernel class.  Its passes re-read each pair's
    coordinates from LDS and L2 many times per launch, so the per-pass bytes
    of the algorithm (24-28 B per aligned pair and pass) exceed what HBM
    delivers (the rocprofv3 FETCH/WRITE counters), and the kernel sits on the
    vector ALU: the roofline is its algorithmic FP32 operations per second
    against the MI355X FP32 vector peak, with the measured HBM rate and the
    counters' VALU-issue fraction beside it."""
    launches = dom["launches"]
    t = dom["seconds"] / launches
    flops = dom["alg_flops"] / launches
    kern = kname.split(" (")[0]
    traffic = profiled_traffic(kern) if full else None
    return {"bound": "valu", "achieved": round(flops / t / 1e12, 3), "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": round(flops / t / FP32_PEAK, 4), "traffic": traffic,
            "hbm_GBps": round(traffic / t / 1e9, 1) if traffic else None,
            "hbm_frac": round(traffic / t / HBM_PEAK, 5) if traffic else None,
            "traffic_source": (TRAFFIC_PROFILE + ": rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per dispatch"
                               if traffic else None),
            "valu_issue_frac": profiled_valu(kern) if full else None,
            "valu_source": VALU_PROFILE if full else None,
            "kernel": kname, "launches": launches, "launch_ms": round(1e3 * t, 3),
            "alg_flops_per_launch": flops, "pass_bytes_per_launch": dom["alg_bytes"] / launches}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_form(qc, rc, prec, steps, warmup, profile, batched=0, tune=None, **opt):
    """Times `steps` calls after `warmup` untimed ones.  With tune (a dict),
    DP launch-form autotuning is on with a fresh cache file: the tuning runs
    inside the first warmup call, and tune receives the seconds it took and
    the pairs per chosen form."""
    opts = default_options(precision=prec, batched=batched, qgroup=QGROUP, **opt)
    with GtaBatch(0, opts, max_qrys=qc.nstrs, max_qry_poss=qc.nposs, max_refs=rc.nstrs,
                  max_ref_poss=rc.nposs) as b:
        b.TransferQueryPMDataToDevice(qc)
        b.TransferCPMDataToDevice(rc)
        if tune is not None:
            import tempfile
            cache = os.path.join(tempfile.mkdtemp(), "dp_forms.txt")
            b.set_autotune(True, cache)
            warmup = max(warmup, 1)
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
        if tune is not None:
            forms, secs = b.dp_forms()
            tune.update({"tuning_seconds": round(secs, 2), "pairs_per_form": forms,
                         "cache_lines": open(cache).read().splitlines() if os.path.exists(cache) else []})
        launched = b.dp_cell_counts()[1] * steps
        stats = None
        if profile:
            b.set_profiling(True)
            s, _ = b.process_resident()
            stats = b.kernel_stats()
            log("  profiled step: %.3f s" % s)
    return wall, dev, (cells, launched), stats


def pcie_inclusive(qc, rc, steps, warmup):
    """The boundary takes host buffers: the plain FP32 form timed over `steps`
    calls with the references' upload (host -> HBM, every call) and the
    results' download (HBM -> host: records, transforms and alignment rows)
    inside the clock, after `warmup` untimed calls on the same context.
    Returns the figures and the last call's results."""
    with GtaBatch(0, default_options(qgroup=QGROUP), max_qrys=qc.nstrs, max_qry_poss=qc.nposs,
                  max_refs=rc.nstrs, max_ref_poss=rc.nposs) as b:
        b.TransferQueryPMDataToDevice(qc)
        b.TransferCPMDataToDevice(rc)
        for _ in range(max(1, warmup)):
            b.ProcessBlock()
        up = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            b.TransferCPMDataToDevice(rc)
            up += time.perf_counter() - t1
            g = b.ProcessBlock()
        wall = time.perf_counter() - t0
    n = qc.nstrs * rc.nstrs
    return {"pairs_per_s": round(n * steps / wall, 1), "steps": steps,
            "upload_s_per_step": round(up / steps, 4),
            "process_and_download_s_per_step": round((wall - up) / steps, 4)}, g


def _ref_run(q, r, nthreads, tmp, fast=False):
    """The reference's CPU path on (q, r) in a child process (its thread count
    is fixed per process); fast selects the reference's own release build
    (-O3 -ffast-math) instead of the deterministic parity build.  Returns
    (alndata, seconds, passed)."""
    import subprocess
    from oracle.ref_cli import save_pair
    tag = "%d%s" % (nthreads, "f" if fast else "")
    src, dst = os.path.join(tmp, "in_%s.npz" % tag), os.path.join(tmp, "out_%s.npz" % tag)
    save_pair(src, q, r)
    env = dict(os.environ, GTA_REF_FAST="1" if fast else "0", OMP_NUM_THREADS=str(nthreads))
    subprocess.run([sys.executable, "-m", "oracle.ref_cli", src, dst, str(nthreads)], cwd=ROOT, check=True,
                   env=env)
    d = np.load(dst, allow_pickle=False)
    return (d["alndata"].reshape(q.nstrs, r.nstrs, -1), float(d["seconds"]),
            d["passed"].reshape(q.nstrs, r.nstrs).astype(bool))


def _gpu_block(q, r, prec=PREC_FP32):
    with GtaBatch(0, default_options(precision=prec, qgroup=QGROUP), max_qrys=q.nstrs, max_qry_poss=q.nposs,
                  max_refs=r.nstrs, max_ref_poss=r.nposs) as b:
        b.TransferQueryPMDataToDevice(q)
        b.TransferCPMDataToDevice(r)
        return b.ProcessBlock()


def _dp_identity(q, r, tfms):
    """Fraction of pairs whose DP alignment under the given transforms is
    identical between the DP kernel and the oracle's DP (GAP0 and gap-open
    -0.6 forms)."""
    from oracle import oracle
    same = total = 0
    for gap0, gapopen in ((True, 0.0), (False, -0.6)):
        with GtaBatch(0, default_options(), max_qrys=q.nstrs, max_qry_poss=q.nposs, max_refs=r.nstrs,
                      max_ref_poss=r.nposs) as b:
            b.TransferQueryPMDataToDevice(q)
            b.TransferCPMDataToDevice(r)
            got = b.dp_under(tfms, gap0, gapopen)
        for qi in range(q.nstrs):
            for ri in range(r.nstrs):
                d02 = float(oracle.d0(int(q.len[qi]), int(r.len[ri]))["d0"]) ** 2
                _, want = oracle.dp(q.coords(qi), r.coords(ri), tfms[qi, ri], d02, gapopen, gap0)
                same += int(np.array_equal(got[qi, ri], want))
                total += 1
    return same / total, total


def _fp64_summary(res, q, r):
    from oracle.fp64 import fp64_errors
    dtm, drm = fp64_errors(res, q, r)
    if not len(dtm):
        return None
    return {"aligned_pairs": int(len(dtm)), "tm_err_max": float(dtm.max()), "tm_err_mean": float(dtm.mean()),
            "rmsd_err_max": float(drm.max()), "rmsd_err_mean": float(drm.mean())}


def _tm_agreement(ref, g, passed=None):
    """TM-score agreement over the pairs the reference reports (those whose
    reference stayed after the stage-1 filter); the GPU's other pairs must
    carry no alignment."""
    if passed is None:
        passed = np.ones(ref.shape[:2], bool)
    aligned = passed & (ref[..., 8] < 9999.0)
    dtm = np.abs(g.alndata[..., 9:11] - ref[..., 9:11]).max(-1)[passed]
    return {"pairs": int(passed.size), "reported_pairs": int(passed.sum()), "aligned_pairs": int(aligned.sum()),
            "same_aligned_set": bool(np.array_equal(aligned, g.alndata[..., 8] < 9999.0)),
            "tm_err_max": float(dtm.max()) if dtm.size else 0.0, "pairs_tm_err_gt_1e-4": int((dtm > 1e-4).sum())}


def _attribute(g, ref_n, pass_n, nthreads):
    """Each pair where the GPU and the reference at `nthreads` threads differ
    (aligned or not, or TM-score by more than 1e-4), attributed with the
    committed 1-thread runs of the same sample (tests/golden/scale_4x8000.npz,
    written by tests/golden/make_scale_golden.py): `thread_count` when the GPU
    equals the 1-thread reference; `packing` when it equals only the 1-thread
    run with per-slot rows moved at packing (the reference's stale per-slot
    rows); `near_tie` when it equals neither and the pair is weak (greater
    TM-score below 0.5 on both sides: a search ending on near-ties between
    superpositions of unrelated fragments); `unexplained` otherwise."""
    from tests.golden.make_scale_golden import FIELDS, OUT
    d = np.load(os.path.join(ROOT, OUT), allow_pickle=False)
    nq, nr = ref_n.shape[:2]
    n = nq * nr
    col = {f: i for i, f in enumerate(FIELDS)}

    def unpack(tag):
        al = np.unpackbits(d[tag + "_aligned"])[:n].reshape(nq, nr).astype(bool)
        tm = np.full((nq, nr, 2), -1.0)
        tm[al] = d[tag + "_values"][:, [col[9], col[10]]]
        return al, tm

    def same(al_a, tm_a, al_b, tm_b):
        return (al_a == al_b) & (~al_a | (np.abs(tm_a - tm_b).max(-1) <= 1e-4))

    g_al = g.alndata[..., 8] < 9999
    g_tm = np.where(g_al[..., None], g.alndata[..., 9:11], -1.0)
    n_al = pass_n & (ref_n[..., 8] < 9999)
    n_tm = np.where(n_al[..., None], ref_n[..., 9:11], -1.0)
    p_al, p_tm = unpack("plain")
    m_al, m_tm = unpack("moved")
    diff = ~same(g_al, g_tm, n_al, n_tm)
    g_eq_p = same(g_al, g_tm, p_al, p_tm)
    g_eq_m = same(g_al, g_tm, m_al, m_tm)
    weak = np.maximum(g_tm.max(-1), p_tm.max(-1)) < 0.5

    def cause(a, b):
        if g_eq_p[a, b]:
            return "thread_count"
        if g_eq_m[a, b]:
            return "packing"
        return "near_tie" if weak[a, b] else "unexplained"

    causes = [cause(a, b) for a, b in np.argwhere(diff)]
    out = {"pairs_differing": int(diff.sum())}
    for c in ("thread_count", "packing", "near_tie", "unexplained"):
        out[c] = causes.count(c)
    out["gpu_vs_1thread"] = {"pairs_differing": int((~g_eq_p).sum()),
                             "weak_pairs_differing": int((~g_eq_p & weak).sum()),
                             "same_aligned_set": bool(np.array_equal(g_al, p_al)),
                             "aligned_pairs": int(p_al.sum())}
    out["reference_%d_vs_1thread_pairs_differing" % nthreads] = int((~same(n_al, n_tm, p_al, p_tm)).sum())
    out["pairs"] = [{"q": int(a), "r": int(b), "cause": c} for (a, b), c in zip(np.argwhere(diff)[:64], causes)]
    return out


def cpu_baseline(qc, rc, src):
    """The reference's own CPU path (oracle/_ref) on a bounded sample of the
    workload (queries 0, 5, 10, 15 x 8,000 references): timed in its own
    release build (-O3 -ffast-math) on every core of this process's affinity
    mask, and run in the deterministic parity build on as many threads for
    the GPU's agreement, each disagreement attributed to the reference's
    thread count or packing with the committed 1-thread runs."""
    import tempfile
    from oracle import oracle
    rng = np.random.default_rng(7)
    qsel = [0, qc.nstrs // 3, (2 * qc.nstrs) // 3, qc.nstrs - 1]
    nref = min(rc.nstrs, 8000)
    rsel = np.sort(rng.choice(rc.nstrs, nref, replace=False))
    q = qc.subset(qsel)
    r = rc.subset(rsel)
    npairs = q.nstrs * r.nstrs
    if not oracle.ref_available():
        q = q.subset([0, 1])
        r = r.subset(list(range(min(400, r.nstrs))))
        t0 = time.perf_counter()
        res = oracle.process_block(q, r)
        dt = time.perf_counter() - t0
        out = {"value": q.nstrs * r.nstrs / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": "%d queries x %d references of the same workload, %.1f s" % (q.nstrs, r.nstrs, dt)}
        out["gpu_vs_port"] = _tm_agreement(res.alndata, _gpu_block(q, r))
        return out
    affinity = sorted(os.sched_getaffinity(0))
    threads = len(affinity)
    fast = os.path.exists(oracle.REF_FAST_LIB)
    with tempfile.TemporaryDirectory() as tmp:
        _, dt, _ = _ref_run(q, r, threads, tmp, fast=fast)
        ref_t, dt_parity, pass_t = _ref_run(q, r, threads, tmp)
        # single-thread parity sample for the DP identity: 2 queries x (their
        # homologs + as many others)
        qs = q.subset([0, 3])
        rsrc = src[rsel]
        hom = np.flatnonzero((rsrc == qsel[0]) | (rsrc == qsel[3]))[:200]
        oth = np.setdiff1d(np.arange(r.nstrs), hom)
        oth = oth[np.linspace(0, len(oth) - 1, len(hom) + 50).astype(int)]
        rs = r.subset(np.union1d(hom, oth))
        ref_1, _, pass_1 = _ref_run(qs, rs, 1, tmp)
    out = {"value": npairs / dt, "unit": "pairs/s", "cores": threads, "kind": "reference",
           "sample": "%d queries x %d references of the same workload (%d pairs), %.1f s with %d threads, "
                     "reference release build (-O3 -ffast-math)%s" %
                     (q.nstrs, r.nstrs, npairs, dt, threads, "" if fast else " absent: parity build timed"),
           "host": {"nproc": os.cpu_count(), "affinity_cpus": threads,
                    "affinity": "%d-%d" % (affinity[0], affinity[-1]) if affinity else ""},
           "parity_build_pairs_per_s": round(npairs / dt_parity, 1)}
    g32, gmx = _gpu_block(qs, rs, PREC_FP32), _gpu_block(qs, rs, PREC_MIXED)
    out["gpu_vs_reference_1thread"] = {"sample": "%d x %d pairs" % (qs.nstrs, rs.nstrs),
                                       "plain": _tm_agreement(ref_1, g32, pass_1),
                                       "mixed": _tm_agreement(ref_1, gmx, pass_1)}
    g = _gpu_block(q, r)
    out["gpu_vs_reference_%dthreads" % threads] = _tm_agreement(ref_t, g, pass_t)
    out["disagreements_attributed"] = _attribute(g, ref_t, pass_t, threads)
    frac, n = _dp_identity(qs, rs, g32.tfms)
    out["dp_bit_identical"] = {"fraction": frac, "fills": n}
    return out


BASE_SEED_ASM = 20261206


def precision_stress(assemblies=True):
    """Near-collinear/near-planar chains against their noisy rigid copies
    (20 queries x 2,000 references, 2,000 copy pairs): max/mean error of the
    reported TM-score and RMSD against FP64 recomputations, plain and mixed."""
    from oracle.fp64 import fp64_errors
    qc, rc, src = synth.config(3)
    out = {"workload": "20 segment chains (40%% straight 3-8 residue segments, rest helical) x %d "
                       "noisy rigid copies (sigma 0.01-0.1 A)" % rc.nstrs}
    for name, prec in (("plain", PREC_FP32), ("mixed", PREC_MIXED)):
        g = _gpu_block(qc, rc, prec)
        # the copy pairs only
        sel = np.zeros((qc.nstrs, rc.nstrs), bool)
        sel[src, np.arange(rc.nstrs)] = True
        g.alndata = np.where(sel[..., None], g.alndata, np.float32(9999.9))
        dtm, drm = fp64_errors(g, qc, rc)
        out[name] = {"copy_pairs_aligned": int(len(dtm)), "tm_err_max": float(dtm.max()),
                     "tm_err_mean": float(dtm.mean()), "rmsd_err_max": float(drm.max()),
                     "rmsd_err_mean": float(drm.mean())}
    if assemblies:
        # two assemblies of 150 domains (37,860 residues), the second a rigid copy
        # of the first with 0.05 A noise: one pair, filled by 16-wave DP, its
        # searches split over workgroups
        from gtalign_alpha_amd.chunk import Chunk
        rng = np.random.default_rng(BASE_SEED_ASM)
        a = synth.assembly(rng)
        b = synth.rigid_copy(rng, a, 0.05)
        q = Chunk.from_structures([a], ["A" * len(a)])
        r = Chunk.from_structures([b], ["A" * len(b)])
        asm = {"residues": int(len(a))}
        for name, prec in (("plain", PREC_FP32), ("mixed", PREC_MIXED)):
            t0 = time.perf_counter()
            g = _gpu_block(q, r, prec)
            dt = time.perf_counter() - t0
            dtm, drm = fp64_errors(g, q, r)
            asm[name] = {"seconds": round(dt, 2), "tm_q": float(g.alndata[0, 0, 9]),
                         "rmsd": float(g.alndata[0, 0, 8]),
                         "tm_err": float(dtm.max()) if len(dtm) else None,
                         "rmsd_err": float(drm.max()) if len(drm) else None}
            log("  assembly pair, %s: %.1f s, %s" % (name, dt, asm[name]))
        out["assemblies"] = asm
    return out


def copy_calibration(mib):
    """One device-to-device copy of mib MiB (the runtime's copy kernel): its
    FETCH_SIZE / WRITE_SIZE counters give the scale of those counters."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


VALU_KINDS = ["v_fma_f32 (compiler-formed, scalar factor)", "v_fma_f32", "v_mul_f32", "v_add_u32", "v_max_f32",
              "v_add_f32 with DPP row shift", "v_exp_f32", "v_mov_b32"]


def valu_calibration(iters, blocks=2048):
    """The VALU calibration kernel alone (blocks x 4 waves; every wave issues
    8 * iters instructions of one kind plus 25 or 26 around its loop): the
    first record is kind 0 on full waves, the instruction count the SQ
    counters are checked against; then the issue rate of every kind on full
    waves and of three kinds with 1 to 48 of 64 lanes active, against the
    FP32 issue ceiling VALU_ISSUE_PEAK."""
    runs = [(k, 64) for k in range(len(VALU_KINDS))]
    runs += [(k, n) for k in (1, 2, 4) for n in (1, 4, 8, 9, 12, 16, 24, 32, 48)]
    out = []
    with GtaBatch(0, default_options(), max_qrys=1, max_qry_poss=1, max_refs=1, max_ref_poss=1) as b:
        for kind, lanes in runs:
            secs = b.valu_calibration(blocks, iters, kind, lanes)
            waves = blocks * 4
            insts = waves * (8 * iters + (26 if 1 <= kind <= 4 else 25))
            out.append({"kind": kind, "instruction": VALU_KINDS[kind], "active_lanes": lanes, "blocks": blocks,
                        "waves": waves, "iters": iters, "valu_wave_instructions": insts, "seconds": secs,
                        "wave_instructions_per_s": insts / secs,
                        "frac_of_issue_peak": round(insts / secs / VALU_ISSUE_PEAK, 4)})
            log("  %-42s %2d lanes: %8.3f ms  %7.1f G wave-instructions/s  %.3f of %.1f G/s" %
                (VALU_KINDS[kind], lanes, secs * 1e3, insts / secs * 1e-9, insts / secs / VALU_ISSUE_PEAK,
                 VALU_ISSUE_PEAK * 1e-9))
    print(json.dumps({"valu_calibration": out[0], "issue_rates": out[1:]}), flush=True)


def batched_vs_plain(steps, warmup):
    """1 query of 400 x 500,000 references (lognormal(220, 0.7) in [20, 4000],
    10% homologs) in one chunk: pairs/s of the plain FP32 form against the
    batched form (device-built live-pair lists, persistent grids, the launch
    sequence as one HIP graph) in FP32 and mixed precision."""
    t0 = time.perf_counter()
    qc, rc, _ = synth.config(2)
    npairs = qc.nstrs * rc.nstrs
    log("batched vs plain: %d x %d pairs (%d + %d residues), generated in %.1f s" %
        (qc.nstrs, rc.nstrs, qc.nposs, rc.nposs, time.perf_counter() - t0))
    out = {"workload": "1 query of 400 x %d references, lognormal(220,0.7) in [20,4000], 10%% homologs"
                       % rc.nstrs, "pairs": npairs}
    for name, prec, batched, tuned in (("plain_fp32", PREC_FP32, 0, False), ("batched_fp32", PREC_FP32, 1, False),
                                       ("batched_mixed", PREC_MIXED, 1, False),
                                       ("batched_mixed_autotuned", PREC_MIXED, 1, True)):
        log(" " + name)
        tune = {} if tuned else None
        wall, dev, (cells, launched), _ = time_form(qc, rc, prec, steps, warmup, False, batched, tune)
        out[name] = {"pairs_per_s": round(npairs * steps / wall, 1), "ms_per_step": round(1e3 * wall / steps, 2),
                     "dp_gcups": round(cells / dev / 1e9, 2), "dp_gcups_launched": round(launched / dev / 1e9, 2)}
        if tuned:
            out[name]["dp_autotune"] = tune
    plain = out["plain_fp32"]["pairs_per_s"]
    out["batched_mixed_vs_plain"] = round(out["batched_mixed"]["pairs_per_s"] / plain, 3)
    out["batched_mixed_autotuned_vs_plain"] = round(out["batched_mixed_autotuned"]["pairs_per_s"] / plain, 3)
    out["batched_fp32_vs_plain"] = round(out["batched_fp32"]["pairs_per_s"] / plain, 3)
    return out


def all_vs_all(steps, warmup):
    """All-vs-all family set without prescreen (--pre-score 0): 16 queries x
    20,000 references of 500 families, batched mixed-precision form; pairs/s,
    GCUPS, the roofline of its dominant kernel class and the FP64 errors of
    every reported TM-score and RMSD."""
    from oracle.fp64 import fp64_errors
    qc, rc, _ = synth.config(4)
    npairs = qc.nstrs * rc.nstrs
    log("all-vs-all: %d x %d pairs, no prescreen" % (qc.nstrs, rc.nstrs))
    opts = default_options(precision=PREC_MIXED, batched=1, prescore=0.0)
    with GtaBatch(0, opts, max_qrys=qc.nstrs, max_qry_poss=qc.nposs, max_refs=rc.nstrs,
                  max_ref_poss=rc.nposs) as b:
        b.TransferQueryPMDataToDevice(qc)
        b.TransferCPMDataToDevice(rc)
        for _ in range(warmup):
            b.process_resident()
        t0 = time.perf_counter()
        dev = cells = 0.0
        for _ in range(steps):
            s_, c_ = b.process_resident()
            dev += s_
            cells += c_
        wall = time.perf_counter() - t0
        b.set_profiling(True)
        b.process_resident()
        stats = b.kernel_stats()
        b.set_profiling(False)
        g = b.ProcessBlock()
    dom = max((st for st in stats if st["launches"]), key=lambda st: st["seconds"])
    dtm, drm = fp64_errors(g, qc, rc)
    fam = np.linspace(0, 499, qc.nstrs).astype(int)
    return {"workload": "16 family seed chains x 20,000 chains of 500 families (lognormal(200,0.5) in "
                        "[40,800], members: rigid motion, noise 0.5-3 A, 0-15% indels); --pre-score 0",
            "tile": "rows %s (generation order: the seed chains of families %s) x all 20,000 columns of the "
                    "20,000 x 20,000 all-vs-all matrix: %d of 20,000 rows, %.2f%% of its pairs" %
                    (",".join(str(40 * f) for f in fam), ",".join(str(f) for f in fam), qc.nstrs,
                     100.0 * qc.nstrs / 20000.0),
            "form": "batched, mixed precision", "pairs": npairs,
            "pairs_per_s": round(npairs * steps / wall, 1), "ms_per_step": round(1e3 * wall / steps, 2),
            "dp_gcups": round(cells / dev / 1e9, 2),
            "roofline": valu_roofline(dom, KERNEL_NAMES.get(dom["name"], dom["name"]), False),
            "aligned_pairs": int(len(dtm)),
            "vs_fp64": {"tm_err_max": float(dtm.max()) if len(dtm) else None,
                        "rmsd_err_max": float(drm.max()) if len(drm) else None}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the 50,000 reference chains (default: the full workload)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--search-by-ss", action="store_true",
                    help="also time the same workload with --add-search-by-ss (off by default so that the "
                         "kernel-trace averages of the default run cover the default options only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stress", action="store_true", help="skip the precision-stress section")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the plain-vs-batched comparison on 1 x 500,000 pairs")
    ap.add_argument("--no-all-vs-all", action="store_true",
                    help="skip the all-vs-all family set (16 x 20,000 pairs, no prescreen)")
    ap.add_argument("--no-assemblies", action="store_true",
                    help="skip the 37,860-residue assembly pair of the precision-stress section")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--valu-calibration", type=int, default=0, metavar="ITERS",
                    help="only run the VALU calibration kernel with ITERS loop iterations (SQ counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.valu_calibration:
        valu_calibration(args.valu_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the stage's pairs shard across GPUs with no exchange "
              "(whole reference chunks per GPU), and runs on more than one GPU are not part of "
              "this build; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    qc, rc, src = synth.config(1, scale=args.scale)
    npairs = qc.nstrs * rc.nstrs
    log("workload: %d queries x %d references = %d pairs (%d + %d residues), generated in %.1f s" %
        (qc.nstrs, rc.nstrs, npairs, qc.nposs, rc.nposs, time.perf_counter() - t0))

    log("plain FP32 form")
    wall, dev, (cells, launched), stats = time_form(qc, rc, PREC_FP32, args.steps, args.warmup, True)
    ms = 1e3 * wall / args.steps
    value = npairs * args.steps / wall
    log("PCIe-inclusive calls (upload + stage + download)")
    pcie, g_full = pcie_inclusive(qc, rc, args.steps, args.warmup)
    log("FP64 recomputation of every aligned pair")
    from oracle.fp64 import fp64_errors_all
    dtm_all, drm_all = fp64_errors_all(g_full, qc, rc)
    del g_full
    mixed = None
    if not args.no_mixed:
        log("mixed-precision form")
        wall_m, dev_m, _, stats_m = time_form(qc, rc, PREC_MIXED, args.steps, args.warmup, True)
        mixed = npairs * args.steps / wall_m
        for st in stats_m:
            if st["launches"]:
                log("  %-16s launches %4d  %8.3f s" % (st["name"], st["launches"], st["seconds"]))

    ssrr = None
    if args.search_by_ss:
        log("plain FP32 form with --add-search-by-ss")
        wall_s, _, cells_s, stats_s = time_form(qc, rc, PREC_FP32, args.steps, args.warmup, True, addsearchbyss=1)
        ssrr = {"pairs_per_s": round(npairs * args.steps / wall_s, 1),
                "ms_per_step": round(1e3 * wall_s / args.steps, 2),
                "vs_default_options": round(wall / wall_s, 3),
                "kernel_seconds": {st["name"]: round(st["seconds"], 4) for st in stats_s if st["launches"]}}

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
                            "alg_GB": round(st["alg_bytes"] / 1e9, 3),
                            "alg_TFLOP": round(st["alg_flops"] / 1e12, 4)}
        if dom is None or st["seconds"] > dom["seconds"]:
            dom = st
    kname = KERNEL_NAMES.get(dom["name"], dom["name"])
    roofline = valu_roofline(dom, kname, args.scale == 1.0)

    batched = None
    if not args.no_batched:
        batched = batched_vs_plain(max(1, min(args.steps, 2)), 1)

    avsa = None
    if not args.no_all_vs_all:
        avsa = all_vs_all(1, 1)

    stress = None
    if not args.no_stress:
        log("precision stress")
        stress = precision_stress(not args.no_assemblies)

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(qc, rc, src)

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
                               "20%% homologs; default options with --no-detailed-search" % rc.nstrs,
                   "pairs": npairs, "form": "plain FP32, one chunk per launch sequence",
                   "query_blocks": QGROUP},
        "dp_gcups": round(cells / dev / 1e9, 2),
        "dp_gcups_launched": round(launched / dev / 1e9, 2),
        "dp_kernel_gcups": round(cells / args.steps / kern["dp"]["seconds"] / 1e9, 2) if "dp" in kern else None,
        "dp_kernel_gcups_launched": (round(launched / args.steps / kern["dp"]["seconds"] / 1e9, 2)
                                     if "dp" in kern else None),
        "vs_fp64_all_aligned": {"aligned_pairs": int(len(dtm_all)),
                                "tm_err_max": float(dtm_all.max()) if len(dtm_all) else None,
                                "tm_err_mean": float(dtm_all.mean()) if len(dtm_all) else None,
                                "pairs_tm_err_gt_1e-4": int((dtm_all > 1e-4).sum()),
                                "rmsd_err_max": float(drm_all.max()) if len(drm_all) else None,
                                "rmsd_err_mean": float(drm_all.mean()) if len(drm_all) else None,
                                "form": "plain FP32 (RMSD from the reference's FP32 E0 identity)"},
        "mixed_precision_pairs_per_s": round(mixed, 1) if mixed else None,
        "search_by_ss": ssrr,
        "pcie_inclusive": pcie,
        "kernels": kern,
        "roofline": roofline,
        "batched_vs_plain": batched,
        "all_vs_all": avsa,
        "precision_stress": stress,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
