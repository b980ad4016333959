"""Run the reference's own CPU path (oracle/_ref) on a chunk pair stored in an
.npz file, in a process of its own (TEST INFRASTRUCTURE ONLY).

The reference fixes its thread count at the first call in a process (function
static CLOptions reads), so runs with different thread counts need separate
processes.  bench.py's CPU baseline uses this.

    python -m oracle.ref_cli IN.npz OUT.npz NTHREADS
IN holds q_len/q_x/q_y/q_z/q_rsd and r_* arrays; OUT receives alndata, tfms,
alns, the per-pair record of the stage-1 filter (passed) and the wall-clock
seconds of the call.
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from gtalign_alpha_amd.chunk import Chunk  # noqa: E402
from oracle import oracle  # noqa: E402


def save_pair(path, q, r):
    np.savez(path, **{p + k: getattr(c, k) for p, c in (("q_", q), ("r_", r))
                      for k in ("len", "x", "y", "z", "rsd")})


def load_pair(path):
    d = np.load(path, allow_pickle=False)

    def chunk(p):
        return Chunk(d[p + "len"], d[p + "x"], d[p + "y"], d[p + "z"], rsd=d[p + "rsd"])

    return chunk("q_"), chunk("r_")


def main():
    src, dst, nthreads = sys.argv[1], sys.argv[2], int(sys.argv[3])
    q, r = load_pair(src)
    t0 = time.perf_counter()
    res = oracle.ref_process_block(q, r, nthreads=nthreads)
    dt = time.perf_counter() - t0
    np.savez(dst, alndata=res.alndata, tfms=res.tfms, alns=res.alns, seconds=dt, passed=res.passed)


if __name__ == "__main__":
    main()
