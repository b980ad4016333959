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
    ref = d["alndata"]
    assert np.array_equal(o.alns, d["alns"])
    for f in (2, 3, 4, 5, 6, 7, 11, 12):
        assert np.array_equal(o.alndata[..., f], ref[..., f]), f
    assert np.abs(o.alndata[..., 9:11] - ref[..., 9:11]).max() < 1e-4
    assert np.abs(o.alndata[..., 8] - ref[..., 8]).max() < 1e-3
    aligned = ref[..., 8] < 9999
    assert np.abs(o.tfms - d["tfms"])[aligned].max() < 1e-3


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 2
