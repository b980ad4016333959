"""Writes the golden fixtures of the stage from the reference's own CPU path.

Run in a tree where oracle/_ref/libgta_ref.so was built from the reference
sources (oracle/Makefile).  Each fixture holds the inputs (chunk arrays of the
seeded synthetic workloads), the reference's three result sections and, per
pair, whether the pair's reference stayed in the block after the stage-1
filter (`passed`; the reference reports nothing for the others).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gtalign_alpha_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402


def chunk_arrays(prefix, c):
    return {prefix + "len": c.len, prefix + "x": c.x, prefix + "y": c.y, prefix + "z": c.z,
            prefix + "rsd": c.rsd}


def write(name, q, r):
    res = oracle.ref_process_block(q, r, nthreads=1)
    kept = res.passed.any(axis=0)
    nk = int(kept.sum())
    assert kept[:nk].all() and not kept[nk:].any(), "kept references must come first"
    d = {}
    d.update(chunk_arrays("q_", q))
    d.update(chunk_arrays("r_", r))
    d["alndata"] = res.alndata
    d["tfms"] = res.tfms
    d["alns"] = res.alns
    d["passed"] = res.passed
    name = "%s_%dx%d" % (name, q.nstrs, r.nstrs)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **d)
    print(path, "pairs", q.nstrs * r.nstrs, "kept", int(res.passed.sum()),
          "aligned", int((res.passed & (res.alndata[..., 8] < 9999)).sum()))


if __name__ == "__main__":
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libgta_ref.so is not built")
    qc, rc, _ = synth.config(0, scale=96 / 256.0)
    write("config1", qc, oracle.kept_first(qc, rc))
    qc, rc, src = synth.config(1, scale=0.01)
    # two queries; homologs of the chosen queries and as many other references
    qsel = [2, 9]
    hom = [i for i in range(rc.nstrs) if src[i] in qsel][:30]
    other = [i for i in range(rc.nstrs) if src[i] < 0][:30]
    q2 = qc.subset(qsel)
    write("config2", q2, oracle.kept_first(q2, rc.subset(sorted(hom + other, key=lambda i: -rc.len[i]))))
