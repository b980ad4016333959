"""The stage without device allocation (gta_process_noalloc) against the
stage that allocates (gta_process): same chunks, same results bit for bit, on
two calls in a row with the second smaller; chunks beyond the scratch are
rejected with GTA_ERR_SIZE and the context keeps working."""
import ctypes as C

import numpy as np
import pytest

from gtalign_alpha_amd import synth
from gtalign_alpha_amd.batch import GtaBatch, lib
from gtalign_alpha_amd.capi import STATUS_CODES, alignment_section_size, default_options

pytestmark = pytest.mark.gpu


def noalloc(b, q, r):
    L = lib()
    L.gta_process_noalloc.argtypes = [C.c_void_p] * 6 + [C.c_size_t]
    L.gta_process_noalloc.restype = C.c_int
    nq, nr = q.nstrs, r.nstrs
    aln = np.full((nq, nr, 15), np.nan, np.float32)
    tfm = np.full((nq, nr, 12), np.nan, np.float32)
    size = alignment_section_size(nq, q.nposs, nr, r.nposs)
    alns = np.full(size, 0xAB, np.uint8)
    qv, rv = q.view(), r.view()
    rc = L.gta_process_noalloc(b._h, C.addressof(qv), C.addressof(rv), aln.ctypes.data, tfm.ctypes.data,
                               alns.ctypes.data, size)
    return rc, aln, tfm, alns


def allocating(b, q, r):
    b.TransferQueryPMDataToDevice(q)
    b.TransferCPMDataToDevice(r)
    return b.ProcessBlock()


def test_noalloc_equals_allocating_entry():
    qc, rc, _ = synth.config(1, scale=0.002)
    q = qc.subset([0, 7, 15])
    big, small = (q, rc), (q.subset([1]), rc.subset(list(range(30, 70))))
    opts = default_options()
    with GtaBatch(0, opts, max_qrys=q.nstrs, max_qry_poss=q.nposs, max_refs=rc.nstrs,
                  max_ref_poss=rc.nposs) as b:
        want = [allocating(b, *big), allocating(b, *small)]
        got = [noalloc(b, *big), noalloc(b, *small)]
        # the context's own buffers still serve gta_process afterwards
        again = allocating(b, *small)
    for w, (rc_, aln, tfm, alns) in zip(want, got):
        assert rc_ == 0, rc_
        assert np.array_equal(aln, w.alndata)
        assert np.array_equal(tfm, w.tfms)
        assert np.array_equal(alns, w.alns)
        assert (aln[..., 8] < 9999).sum() >= 3
    assert np.array_equal(again.alndata, want[1].alndata)


def test_noalloc_rejects_chunks_beyond_scratch():
    qc, rc, _ = synth.config(1, scale=0.05)          # 16 x 2,500 pairs: > 64 MiB of pair state
    small_q, small_r = qc.subset([3]), rc.subset(list(range(100, 130)))
    with GtaBatch(0, default_options(), max_qrys=qc.nstrs, max_qry_poss=qc.nposs, max_refs=rc.nstrs,
                  max_ref_poss=rc.nposs) as b:
        rc_, aln, _, _ = noalloc(b, qc, rc)
        assert rc_ == STATUS_CODES["GTA_ERR_SIZE"], rc_
        assert np.isnan(aln).all()                     # nothing written
        rc2, aln2, tfm2, alns2 = noalloc(b, small_q, small_r)
        assert rc2 == 0
        w = allocating(b, small_q, small_r)
    assert np.array_equal(aln2, w.alndata)
    assert np.array_equal(tfm2, w.tfms)
