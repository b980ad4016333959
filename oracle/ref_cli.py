"""Run the reference's own CPU path (oracle/_ref) on a chunk pair stored in an
.npz file, in a process of its own (TEST INFRASTRUCTURE ONLY).

The reference fixes its thread count at the first call in a process (function
static CLOptions reads), so runs with different thread counts need separate
processes.  bench.py's CPU baseline uses this.

    python -m oracle.ref_cli IN.npz OUT.npz NTHREADS [OPTIONS_JSON]
IN holds q_len/q_x/q_y/q_z/q_rsd (and optionally q_ins_ch_ord) and r_*
arrays; OUT receives alndata, tfms, alns, the per-pair record of the stage-1
filter (passed) and the wall-clock seconds of the call.  OPTIONS_JSON maps
gta_options field names to values (the reference defaults otherwise); the
reference reads its options once per process, so each option set needs a
process of its own.
"""
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from gtalign_alpha_amd.chunk import Chunk  # noqa: E402
from oracle import oracle  # noqa: E402


def save_pair(path, q, r):
    np.savez(path, **{p + k: getattr(c, k) for p, c in (("q_", q), ("r_", r))
                      for k in ("len", "x", "y", "z", "rsd", "ins_ch_ord")})


def load_pair(path):
    d = np.load(path, allow_pickle=False)

    def chunk(p):
        ico = d[p + "ins_ch_ord"] if (p + "ins_ch_ord") in d else None
        return Chunk(d[p + "len"], d[p + "x"], d[p + "y"], d[p + "z"], rsd=d[p + "rsd"], ins_ch_ord=ico)

    return chunk("q_"), chunk("r_")


def run(q, r, nthreads=1, opts=None, tmpdir=None):
    """The reference's CPU path on (q, r) in a child process; returns the
    dict of OUT's arrays."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory(dir=tmpdir) as td:
        src, dst = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        save_pair(src, q, r)
        cmd = [sys.executable, "-m", "oracle.ref_cli", src, dst, str(nthreads)]
        if opts:
            cmd.append(json.dumps(opts))
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        subprocess.run(cmd, check=True, cwd=root)
        with np.load(dst, allow_pickle=False) as d:
            return {k: d[k] for k in d.files}


def main():
    src, dst, nthreads = sys.argv[1], sys.argv[2], int(sys.argv[3])
    from gtalign_alpha_amd.capi import default_options
    opts = default_options(**(json.loads(sys.argv[4]) if len(sys.argv) > 4 else {}))
    q, r = load_pair(src)
    t0 = time.perf_counter()
    res = oracle.ref_process_block(q, r, opts, nthreads=nthreads)
    dt = time.perf_counter() - t0
    np.savez(dst, alndata=res.alndata, tfms=res.tfms, alns=res.alns, seconds=dt, passed=res.passed)


if __name__ == "__main__":
    main()
