"""The oracle against the golden fixtures written by the reference's own CPU
path (tests/golden/make_golden.py); runs without the reference tree."""
import glob
import os

import numpy as np
import pytest

from gtalign_alpha_amd.chunk import Chunk
from oracle import oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    d = np.load(path, allow_pickle=False)

    def chunk(p):
        lens = d[p + "len"]
        return Chunk(lens, d[p + "x"], d[p + "y"], d[p + "z"], rsd=d[p + "rsd"])

    return chunk("q_"), chunk("r_"), d


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    q, r, d = load(path)
    o = oracle.process_block(q, r)
    # bit for bit: records, transforms and alignment rows
    assert np.array_equal(o.alns, d["alns"])
    assert np.array_equal(o.alndata, d["alndata"])
    assert np.array_equal(o.tfms, d["tfms"])


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 2
