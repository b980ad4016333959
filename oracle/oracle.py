"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

libgta_oracle.so is the CPU restatement of the reference path
(oracle/gta_oracle.cpp); _ref/libgta_ref.so is the reference's own CPU path
compiled from its sources (oracle/Makefile), present only where the
reference tree was available at build time.  Only tests/, the smoke check
and bench.py's CPU baseline use this module.
"""
import ctypes as C
import os

import numpy as np

from gtalign_alpha_amd.capi import (ChunkView, Options, GTA_ALN_NFIELDS, GTA_TFM_NFIELDS,
                                    alignment_section_size, default_options)

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libgta_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgta_ref.so")
# the reference's own release build (-O3 -ffast-math): CPU timing baseline only
REF_FAST_LIB = os.path.join(HERE, "_ref", "libgta_ref_fast.so")
if os.environ.get("GTA_REF_FAST") == "1":
    REF_LIB = REF_FAST_LIB

_orc = None
_ref = None
_fp = C.POINTER(C.c_float)


def lib():
    global _orc
    if _orc is None:
        _orc = C.CDLL(ORACLE_LIB)
        _orc.gta_oracle_process_block.argtypes = [C.POINTER(ChunkView), C.POINTER(ChunkView),
                                                  C.POINTER(Options), C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _orc.gta_oracle_process_block.restype = C.c_double
        _orc.gta_oracle_dp.argtypes = ([C.c_void_p] * 4 + [C.c_int] + [C.c_void_p] * 4 +
                                       [C.c_int, C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_int,
                                        C.c_float, C.c_void_p, C.c_void_p, C.c_void_p])
        _orc.gta_oracle_dp.restype = C.c_int
        _orc.gta_oracle_dp_score.argtypes = ([C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3 +
                                             [C.c_int, C.c_void_p, C.c_float, C.c_int])
        _orc.gta_oracle_dp_score.restype = C.c_float
        _orc.gta_oracle_dp_ssrr.argtypes = ([C.c_char_p, C.c_char_p, C.c_int] * 2 + [C.c_float] * 3 +
                                            [C.c_void_p] * 3)
        _orc.gta_oracle_dp_ssrr.restype = C.c_int
        _orc.gta_oracle_gonnet.argtypes = [C.c_void_p]
        _orc.gta_oracle_kabsch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _orc.gta_oracle_rmsd.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        _orc.gta_oracle_rmsd.restype = C.c_float
        _orc.gta_oracle_d0.argtypes = [C.c_int, C.c_int, C.c_void_p]
        _orc.gta_oracle_fraglen.argtypes = [C.c_int] * 5
        _orc.gta_oracle_fraglen.restype = C.c_int
        _orc.gta_oracle_secstr.argtypes = [C.POINTER(ChunkView), C.c_void_p]
    return _orc


def ref_available():
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        _ref.gta_ref_process_block.argtypes = [C.POINTER(ChunkView), C.POINTER(ChunkView),
                                               C.POINTER(Options), C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_size_t, C.POINTER(C.c_double)]
        _ref.gta_ref_process_block.restype = C.c_int
        _ref.gta_ref_last_pass.argtypes = [C.c_void_p, C.c_size_t]
        _ref.gta_ref_last_pass.restype = C.c_int
        _ref.gta_ref_set_move_slots.argtypes = [C.c_int]
        _ref.gta_ref_set_move_slots.restype = C.c_int
    return _ref


class BlockResult:
    """Result sections of a block; `passed[q, r]` is True for the pairs whose
    reference stayed in its block after the stage-1 filter (the reference
    reports nothing for the others)."""

    def __init__(self, nq, nr, alndata, tfms, alns, seconds=0.0, cells=0.0, ss_q=None, ss_r=None,
                 passed=None):
        self.nq, self.nr = nq, nr
        self.passed = (np.ones((nq, nr), bool) if passed is None
                       else np.asarray(passed, bool).reshape(nq, nr))
        self.alndata = alndata.reshape(nq, nr, GTA_ALN_NFIELDS)
        self.tfms = tfms.reshape(nq, nr, GTA_TFM_NFIELDS)
        self.alns = alns
        self.seconds = seconds
        self.cells = cells
        self.ss_q, self.ss_r = ss_q, ss_r


def _buffers(q, r):
    nq, nr = q.nstrs, r.nstrs
    aln = np.zeros(nq * nr * GTA_ALN_NFIELDS, dtype=np.float32)
    tfm = np.zeros(nq * nr * GTA_TFM_NFIELDS, dtype=np.float32)
    alns = np.zeros(alignment_section_size(nq, q.nposs, nr, r.nposs), dtype=np.uint8)
    return aln, tfm, alns


def process_block(q, r, opts=None):
    """The whole stage on the oracle."""
    opts = opts or default_options()
    aln, tfm, alns = _buffers(q, r)
    ssq = np.zeros(q.nposs, dtype=np.uint8)
    ssr = np.zeros(r.nposs, dtype=np.uint8)
    passed = np.zeros(q.nstrs * r.nstrs, dtype=np.uint8)
    import time
    t0 = time.perf_counter()
    cells = lib().gta_oracle_process_block(C.byref(q.view()), C.byref(r.view()), C.byref(opts),
                                           aln.ctypes.data, tfm.ctypes.data, alns.ctypes.data,
                                           ssq.ctypes.data, ssr.ctypes.data, passed.ctypes.data)
    dt = time.perf_counter() - t0
    return BlockResult(q.nstrs, r.nstrs, aln, tfm, alns, dt, cells, ssq, ssr, passed)


def last_prod_negative(nq, nr):
    """Per pair of the last process_block call: True when the production full
    search (--refinement FULL/FULLA) had slots at negative alignment
    positions (run by the reference on memory past the alignment, skipped
    here)."""
    f = lib().gta_oracle_last_prod_negative
    f.argtypes = [C.c_void_p, C.c_size_t]
    f.restype = C.c_int
    out = np.zeros(nq * nr, np.uint8)
    if f(out.ctypes.data, out.size) != 0:
        raise RuntimeError("no process_block record")
    return out.reshape(nq, nr).astype(bool)


def set_na_ss_mode(mode):
    """Nucleic-acid pairing partner the oracle records: 0 the stage's
    definition (least deviation, lowest index on ties), 1 the reference CPU
    twin as compiled (its 4-lane vectorised minimum keeps the index of the
    last lane improvement).  Returns the previous mode."""
    f = lib().gta_oracle_set_na_ss_mode
    f.argtypes = [C.c_int]
    f.restype = C.c_int
    return f(int(mode))


def secstr(c):
    """The oracle's secondary structure of every position of chunk c (bytes)."""
    out = np.zeros(c.nposs, dtype=np.uint8)
    lib().gta_oracle_secstr(C.byref(c.view()), out.ctypes.data)
    return out


def ref_secstr(c, with_pairs=False):
    """The reference's own MpSecStr on chunk c (oracle/_ref); with_pairs also
    returns the pairing partner it recorded per position (-1: none)."""
    f = ref_lib().gta_ref_secstr
    f.argtypes = [C.POINTER(ChunkView), C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    out = np.zeros(c.nposs, dtype=np.uint8)
    pos = np.zeros(c.nposs, dtype=np.int32)
    if f(C.byref(c.view()), out.ctypes.data, pos.ctypes.data) != 0:
        raise RuntimeError("reference SS assignment failed")
    return (out, pos) if with_pairs else out


def ref_process_block(q, r, opts=None, nthreads=1):
    """The whole stage on the reference's own CPU code (oracle/_ref)."""
    opts = opts or default_options()
    aln, tfm, alns = _buffers(q, r)
    secs = C.c_double(0)
    rc = ref_lib().gta_ref_process_block(C.byref(q.view()), C.byref(r.view()), C.byref(opts),
                                         nthreads, aln.ctypes.data, tfm.ctypes.data,
                                         alns.ctypes.data, alns.nbytes, C.byref(secs))
    if rc != 0:
        raise RuntimeError("reference CPU path failed (%d)" % rc)
    passed = np.zeros(q.nstrs * r.nstrs, dtype=np.uint8)
    if ref_lib().gta_ref_last_pass(passed.ctypes.data, passed.size) != 0:
        raise RuntimeError("reference CPU path: no filter record")
    return BlockResult(q.nstrs, r.nstrs, aln, tfm, alns, secs.value, passed=passed)


def ref_set_move_slots(on):
    """Diagnostic switch of the reference driver: when on, packing a block
    after stage 1 also moves every kept pair's per-slot transforms and
    records to its new index (the reference moves only the per-pair records
    and output transform).  Returns the previous setting."""
    return ref_lib().gta_ref_set_move_slots(int(on))


def kept_first(q, r, opts=None):
    """The references of r reordered so that, after the reference's own
    filter after stage 1 (one block: all queries of q), the references it
    keeps come before those it drops: kept references longer than a length
    split and dropped ones not longer, in descending length, the split chosen
    to keep the most aligned references.  Kept references then keep their
    indices when the reference packs the block, which leaves fewer per-slot
    buffers behind at a kept pair's index (packing moves the per-pair records
    but not the per-slot buffers, MpReform.cpp:57-124)."""
    res = ref_process_block(q, r, opts, nthreads=1)
    kept = res.passed.any(axis=0)
    aligned = (res.passed & (res.alndata[..., 8] < 9999)).any(axis=0)

    def pick(x):
        return [i for i in range(r.nstrs) if (kept[i] and r.len[i] > x) or (not kept[i] and r.len[i] <= x)]

    x = max(sorted(set(r.len.tolist())), key=lambda v: (sum(aligned[i] for i in pick(v)), len(pick(v))))
    return r.subset(sorted(pick(x), key=lambda i: -r.len[i]))


def dp(qxyz, rxyz, tfm, d02, gapopen=0.0, gap0=True, qss=None, rss=None, ssw=0.0):
    """DP fill + backtrace for one pair; returns (btck[Lq,Lr], matched pairs in walk order)."""
    qxyz = np.ascontiguousarray(qxyz, np.float32)
    rxyz = np.ascontiguousarray(rxyz, np.float32)
    Lq, Lr = len(qxyz), len(rxyz)
    qc = [np.ascontiguousarray(qxyz[:, i]) for i in range(3)]
    rc = [np.ascontiguousarray(rxyz[:, i]) for i in range(3)]
    usess = qss is not None
    qs = np.frombuffer(qss.encode(), np.uint8).copy() if usess else np.zeros(Lq, np.uint8)
    rs = np.frombuffer(rss.encode(), np.uint8).copy() if usess else np.zeros(Lr, np.uint8)
    bt = np.zeros((Lq, Lr), dtype=np.uint8)
    qp = np.zeros(Lq + Lr, dtype=np.int32)
    rp = np.zeros(Lq + Lr, dtype=np.int32)
    t = np.ascontiguousarray(tfm, np.float32)
    n = lib().gta_oracle_dp(qc[0].ctypes.data, qc[1].ctypes.data, qc[2].ctypes.data, qs.ctypes.data,
                            Lq, rc[0].ctypes.data, rc[1].ctypes.data, rc[2].ctypes.data,
                            rs.ctypes.data, Lr, t.ctypes.data, d02, gapopen, int(gap0), int(usess),
                            ssw, bt.ctypes.data, qp.ctypes.data, rp.ctypes.data)
    return bt, np.stack([qp[:n], rp[:n]], axis=1)


def _ssrr_call(fn, qss, qrsd, rss, rrsd, gapopen, wss, wrr):
    Lq, Lr = len(qss), len(rss)
    assert len(qrsd) == Lq and len(rrsd) == Lr
    bt = np.zeros((Lq, Lr), np.uint8)
    qpos = np.zeros(Lq + Lr, np.int32)
    rpos = np.zeros(Lq + Lr, np.int32)
    n = fn(qss, qrsd, Lq, rss, rrsd, Lr, gapopen, wss, wrr, bt.ctypes.data, qpos.ctypes.data,
           rpos.ctypes.data)
    if n < 0:
        raise RuntimeError("DP failed (%d)" % n)
    return bt, np.stack([qpos[:n], rpos[:n]], 1)


def dp_ssrr(qss, qrsd, rss, rrsd, gapopen=-1.0, wss=1.0, wrr=0.2):
    """The SS/residue DP of the --add-search-by-ss search for one pair (bytes
    of SS letters and residue codes); returns (btck[Lq,Lr], matched pairs in
    walk order)."""
    return _ssrr_call(lib().gta_oracle_dp_ssrr, qss, qrsd, rss, rrsd, gapopen, wss, wrr)


def ref_dp_ssrr(qss, qrsd, rss, rrsd, gapopen=-1.0, wss=1.0, wrr=0.2):
    """The same on the reference's own code (oracle/_ref)."""
    f = ref_lib().gta_ref_dp_ssrr
    f.argtypes = [C.c_char_p, C.c_char_p, C.c_int] * 2 + [C.c_float] * 3 + [C.c_void_p] * 3
    f.restype = C.c_int
    return _ssrr_call(f, qss, qrsd, rss, rrsd, gapopen, wss, wrr)


def gonnet():
    """The residue score table (26 x 26, bits) of the oracle."""
    t = np.zeros(26 * 26, np.float32)
    lib().gta_oracle_gonnet(t.ctypes.data)
    return t.reshape(26, 26)


def ref_gonnet():
    """The reference's own table (GONNET_SCORES)."""
    t = np.zeros(26 * 26, np.float32)
    ref_lib().gta_ref_gonnet(C.c_void_p(t.ctypes.data))
    return t.reshape(26, 26)


def dp_score(qxyz, rxyz, tfm, gap0=True, gapopen=0.0):
    """Score-only DP fill (ExecDPScore128xKernel, MpDPHub.cpp:389-647) of one
    pair under tfm with the search d0: the bottom-right cell's value."""
    qxyz = np.ascontiguousarray(qxyz, np.float32)
    rxyz = np.ascontiguousarray(rxyz, np.float32)
    qc = [np.ascontiguousarray(qxyz[:, i]) for i in range(3)]
    rc = [np.ascontiguousarray(rxyz[:, i]) for i in range(3)]
    t = np.ascontiguousarray(tfm, np.float32)
    return float(lib().gta_oracle_dp_score(qc[0].ctypes.data, qc[1].ctypes.data, qc[2].ctypes.data, len(qxyz),
                                           rc[0].ctypes.data, rc[1].ctypes.data, rc[2].ctypes.data, len(rxyz),
                                           t.ctypes.data, gapopen, int(gap0)))


def kabsch(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    t = np.zeros(12, np.float32)
    lib().gta_oracle_kabsch(x.ctypes.data, y.ctypes.data, len(x), t.ctypes.data)
    return t


def rmsd(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return float(lib().gta_oracle_rmsd(x.ctypes.data, y.ctypes.data, len(x)))


def d0(qlen, rlen):
    out = np.zeros(5, np.float32)
    lib().gta_oracle_d0(qlen, rlen, out.ctypes.data)
    return dict(d0=out[0], d02_dpscan=out[1], d0fin=out[2], d82=out[3], d0s=out[4])


def fraglen(qlen, rlen, qpos, rpos, ndx):
    return lib().gta_oracle_fraglen(qlen, rlen, qpos, rpos, ndx)


def ref_dp_score(qxyz, rxyz, tfms):
    """The reference's own score-only DP (oracle/_ref, ExecDPScore128xKernel
    with GAP0) of one pair under K transforms tfms[K, 12]: K scores."""
    L = ref_lib()
    if not hasattr(L, "_dpscore_ready"):
        L.gta_ref_dpscore.argtypes = ([C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3 +
                                      [C.c_int, C.c_void_p, C.c_int, C.c_void_p])
        L.gta_ref_dpscore.restype = C.c_int
        L._dpscore_ready = True
    qxyz = np.ascontiguousarray(qxyz, np.float32)
    rxyz = np.ascontiguousarray(rxyz, np.float32)
    qc = [np.ascontiguousarray(qxyz[:, i]) for i in range(3)]
    rc = [np.ascontiguousarray(rxyz[:, i]) for i in range(3)]
    t = np.ascontiguousarray(tfms, np.float32).reshape(-1, 12)
    out = np.zeros(len(t), np.float32)
    rc_ = L.gta_ref_dpscore(qc[0].ctypes.data, qc[1].ctypes.data, qc[2].ctypes.data, len(qxyz), rc[0].ctypes.data,
                            rc[1].ctypes.data, rc[2].ctypes.data, len(rxyz), t.ctypes.data, len(t),
                            out.ctypes.data)
    if rc_ < 0:
        raise RuntimeError("reference score-only DP failed (%d)" % rc_)
    return out


def ref_dp(qxyz, rxyz, tfm, gap0=True, gapopen=0.0):
    """The reference's own DP (oracle/_ref) for one pair under tfm:
    returns (btck[Lq,Lr], matched pairs in walk order)."""
    L = ref_lib()
    if not hasattr(L, "_dp_ready"):
        L.gta_ref_dp.argtypes = ([C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3 +
                                 [C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_void_p,
                                  C.c_void_p])
        L.gta_ref_dp.restype = C.c_int
        L._dp_ready = True
    qxyz = np.ascontiguousarray(qxyz, np.float32)
    rxyz = np.ascontiguousarray(rxyz, np.float32)
    Lq, Lr = len(qxyz), len(rxyz)
    qc = [np.ascontiguousarray(qxyz[:, i]) for i in range(3)]
    rc = [np.ascontiguousarray(rxyz[:, i]) for i in range(3)]
    bt = np.zeros((Lq, Lr), dtype=np.uint8)
    qp = np.zeros(Lq + Lr, dtype=np.int32)
    rp = np.zeros(Lq + Lr, dtype=np.int32)
    t = np.ascontiguousarray(tfm, np.float32)
    n = L.gta_ref_dp(qc[0].ctypes.data, qc[1].ctypes.data, qc[2].ctypes.data, Lq, rc[0].ctypes.data,
                     rc[1].ctypes.data, rc[2].ctypes.data, Lr, t.ctypes.data, int(gap0), gapopen,
                     bt.ctypes.data, qp.ctypes.data, rp.ctypes.data)
    if n < 0:
        raise RuntimeError("reference DP failed (%d)" % n)
    return bt, np.stack([qp[:n], rp[:n]], axis=1)
