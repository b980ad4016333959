"""Writes the golden fixtures of the stage from the reference's own CPU path.

Run in a tree where oracle/_ref/libgta_ref.so was built from the reference
sources (oracle/Makefile).  Each fixture holds the inputs (chunk arrays of the
seeded synthetic workloads) and the reference's three result sections.
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
    d = {}
    d.update(chunk_arrays("q_", q))
    d.update(chunk_arrays("r_", r))
    d["alndata"] = res.alndata
    d["tfms"] = res.tfms
    d["alns"] = res.alns
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **d)
    print(path, "pairs", q.nstrs * r.nstrs, "aligned", int((res.alndata[..., 8] < 9999).sum()))


if __name__ == "__main__":
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libgta_ref.so is not built")
    qc, rc, _ = synth.config(0, scale=48 / 256.0)
    write("config1_48refs", qc, rc)
    qc, rc, src = synth.config(1, scale=0.01)
    # two queries; 40 references, 20 of them homologs of the chosen queries
    qsel = [2, 9]
    hom = [i for i in range(rc.nstrs) if src[i] in qsel][:20]
    other = [i for i in range(rc.nstrs) if src[i] < 0][:40 - len(hom)]
    write("config2_2x40", qc.subset(qsel), rc.subset(sorted(hom + other, key=lambda i: -rc.len[i])))
