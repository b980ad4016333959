"""The oracle against the golden fixtures written by the reference's own CPU
path (tests/golden/make_golden.py); runs without the reference tree."""
import glob
import os

import numpy as np
import pytest

from gtalign_alpha_amd.chunk import Chunk
from oracle import oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "config*_*.npz")))


def load(path):
    d = np.load(path, allow_pickle=False)

    def chunk(p):
        lens = d[p + "len"]
        return Chunk(lens, d[p + "x"], d[p + "y"], d[p + "z"], rsd=d[p + "rsd"])

    return chunk("q_"), chunk("r_"), d


def check_golden(got, d):
    """Bit for bit against a fixture: the same pairs stay after the stage-1
    filter; for those, records, transforms and alignment rows are equal; the
    pairs of dropped references (which the reference does not report) carry
    no alignment."""
    passed = d["passed"].astype(bool)
    assert np.array_equal(got.passed, passed)
    assert np.array_equal(got.alns, d["alns"])
    assert np.array_equal(got.alndata[passed], d["alndata"][passed])
    assert np.array_equal(got.tfms[passed], d["tfms"][passed])
    assert (got.alndata[~passed][..., 8] >= 9999).all()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    q, r, d = load(path)
    check_golden(oracle.process_block(q, r), d)


def test_golden_fixtures_present():
    """Both fixtures, one with references dropped by the stage-1 filter and
    one whose two queries share a block."""
    assert len(GOLDEN) >= 2
    dropped = [not load(p)[2]["passed"].all() for p in GOLDEN]
    assert any(dropped)
    assert any(load(p)[0].nstrs == 2 for p in GOLDEN)
