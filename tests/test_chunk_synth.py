"""Chunk format and synthetic workload invariants."""
import numpy as np
import pytest

from gtalign_alpha_amd import synth
from gtalign_alpha_amd.chunk import Chunk, MAX_STRUCTURE_LENGTH
from oracle.fp64 import kabsch_svd


def test_chunk_sorted_and_prefix_offsets():
    qc, rc, _ = synth.config(1, scale=0.002)
    for c in (qc, rc):
        assert np.all(np.diff(c.len) <= 0)
        assert c.dst[0] == 0
        assert np.array_equal(c.dst[1:], np.cumsum(c.len[:-1]))
        assert c.nposs == len(c.x) == len(c.rsd)


def test_generator_is_deterministic():
    a = synth.chain(42, 100)
    b = synth.chain(42, 100)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, synth.chain(43, 100))


def test_chain_geometry():
    c = synth.chain(7, 300)
    bonds = np.linalg.norm(np.diff(c, axis=0), axis=1)
    assert np.all(bonds > 2.5) and np.all(bonds < 5.0)
    assert np.linalg.norm(c, axis=1).max() <= 2.2 * 300 ** 0.38 + 1e-3


def test_homolog_is_a_noisy_rigid_copy():
    src = synth.chain(9, 200)
    h, idx = synth.homolog(10, src, 0.5, 40.0, 0.0, 0.0)
    assert np.array_equal(idx, np.arange(200))
    R, t = kabsch_svd(src, h)
    rmsd = np.sqrt(((src @ R.T + t - h) ** 2).sum(1).mean())
    assert 0.3 < rmsd < 1.2


def test_structure_length_limit():
    with pytest.raises(ValueError):
        Chunk.from_structures([np.zeros((MAX_STRUCTURE_LENGTH + 1, 3))])
