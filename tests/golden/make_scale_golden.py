"""Writes the at-scale agreement fixture from the reference's own CPU path.

The sample is the one bench.py's CPU baseline times: of the 16 x 50,000-pair
workload (lognormal(250, 0.5) in [30, 1000], 20% homologs), queries 0, 5, 10
and 15 against 8,000 references drawn with numpy's default_rng(7) -- 32,000
pairs.  The reference runs at 1 thread, in blocks of 2 queries as its
scheduler forms them, twice:
- `moved`: with the reference driver's switch that moves every kept pair's
  per-slot transforms and records along when a block is packed after the
  stage-1 filter (the reference leaves them at the old index,
  MpReform.cpp:57-124, MpStageBase.h:1285-1305);
- `plain`: as the reference runs.
Stored per run: the filter record (`passed`), the aligned-pair mask and, for
the aligned pairs, RMSD, TM-score (query and reference normalisation) and the
integer statistics; plus a checksum of the sample's inputs.  A pair where the
two runs differ owes the difference to the stale per-slot rows of packing.

Run in a tree where oracle/_ref/libgta_ref.so was built (oracle/Makefile).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gtalign_alpha_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(HERE, "scale_4x8000.npz")
FIELDS = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12]


def sample():
    """The CPU-baseline sample of bench.py (same selection)."""
    qc, rc, src = synth.config(1)
    rng = np.random.default_rng(7)
    qsel = [0, qc.nstrs // 3, (2 * qc.nstrs) // 3, qc.nstrs - 1]
    rsel = np.sort(rng.choice(rc.nstrs, min(rc.nstrs, 8000), replace=False))
    return qc.subset(qsel), rc.subset(rsel)


def checksum(q, r):
    return np.array([float(np.sum(c.x, dtype=np.float64) + 2 * np.sum(c.y, dtype=np.float64) +
                           3 * np.sum(c.z, dtype=np.float64)) for c in (q, r)] +
                    [float(q.len.sum()), float(r.len.sum()), float(q.nstrs), float(r.nstrs)])


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libgta_ref.so is not built")
    from gtalign_alpha_amd.capi import default_options
    q, r = sample()
    d = {"checksum": checksum(q, r), "fields": np.array(FIELDS)}
    for tag, move in (("moved", 1), ("plain", 0)):
        oracle.ref_set_move_slots(move)
        res = oracle.ref_process_block(q, r, default_options(qgroup=2), nthreads=1)
        passed = res.passed.astype(bool)
        aligned = passed & (res.alndata[..., 8] < 9999)
        d[tag + "_passed"] = np.packbits(passed.reshape(-1))
        d[tag + "_aligned"] = np.packbits(aligned.reshape(-1))
        d[tag + "_values"] = res.alndata[aligned][:, FIELDS].astype(np.float32)
        print(tag, "reported", int(passed.sum()), "aligned", int(aligned.sum()), "%.1f s" % res.seconds,
              flush=True)
    oracle.ref_set_move_slots(0)
    np.savez_compressed(OUT, **d)
    print(OUT)


if __name__ == "__main__":
    main()
