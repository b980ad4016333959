"""Host-side mirror of the reference's GPU batch object, over the C ABI.

CuBatch (src/libmycu/cubatch/CuBatch.cuh:38-241) is constructed per device,
receives a query chunk (TransferQueryPMDataToDevice, CuBatch.cuh:54-62) and a
reference chunk (TransferCPMDataToDevice, CuBatch.cuh:64-72), and
ProcessBlock (CuBatch.cuh:74-85) computes per pair TM-scores, RMSD, the
superposition and the alignment.  GtaBatch keeps those names and that call
order; errors from the library raise GtaError (the reference throws
myruntime_error).  The HIP library is required: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .capi import (ChunkView, Options, Results, STATUS, GTA_ALN_NFIELDS, GTA_TFM_NFIELDS,
                   default_options)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgta_amd.so")
_lib = None


class _KStat(C.Structure):
    _fields_ = [("name", C.c_char_p), ("launches", C.c_int), ("seconds", C.c_double),
                ("alg_bytes", C.c_double), ("alg_flops", C.c_double)]


class GtaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


def lib():
    """The HIP library (fails loudly when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built; run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.gta_default_options.argtypes = [C.POINTER(Options)]
        L.gta_create.argtypes = [C.c_int, C.POINTER(Options), C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(C.c_void_p)]
        L.gta_create.restype = C.c_int
        L.gta_destroy.argtypes = [C.c_void_p]
        L.gta_set_queries.argtypes = [C.c_void_p, C.POINTER(ChunkView)]
        L.gta_set_queries.restype = C.c_int
        L.gta_set_references.argtypes = [C.c_void_p, C.POINTER(ChunkView)]
        L.gta_set_references.restype = C.c_int
        L.gta_process.argtypes = [C.c_void_p, C.POINTER(Results)]
        L.gta_process.restype = C.c_int
        L.gta_process_resident.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gta_process_resident.restype = C.c_int
        L.gta_last_error.argtypes = [C.c_void_p]
        L.gta_last_error.restype = C.c_char_p
        L.gta_alignment_section_size.argtypes = [C.c_int] * 4
        L.gta_alignment_section_size.restype = C.c_size_t
        L.gta_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.gta_set_profiling.restype = C.c_int
        L.gta_kernel_stats.argtypes = [C.c_void_p, C.POINTER(_KStat), C.c_int]
        L.gta_kernel_stats.restype = C.c_int
        L.gta_dp_cell_counts.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gta_dp_cell_counts.restype = C.c_int
        L.gta_test_valu_calibration.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.POINTER(C.c_double)]
        L.gta_test_valu_calibration.restype = C.c_int
        L.gta_test_dp.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p]
        L.gta_test_dp.restype = C.c_int
        L.gta_test_set_dp_buckets.argtypes = [C.c_void_p, C.c_longlong, C.c_int]
        L.gta_test_set_dp_buckets.restype = C.c_int
        L.gta_test_dp_score.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                                         C.c_void_p]
        L.gta_test_dp_score.restype = C.c_int
        L.gta_set_autotune.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        L.gta_set_autotune.restype = C.c_int
        L.gta_test_dp_forms.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.gta_test_dp_forms.restype = C.c_int
        L.gta_test_set_dp_grid.argtypes = [C.c_void_p, C.c_int]
        L.gta_test_set_dp_grid.restype = C.c_int
        L.gta_test_set_split.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.gta_test_set_split.restype = C.c_int
        L.gta_test_pair_score.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.gta_test_pair_score.restype = C.c_int
        _lib = L
    return _lib


class BlockResults:
    """Host copy of the three result sections of one ProcessBlock."""

    def __init__(self, res):
        nq, nr = res.nqrys, res.nrefs
        self.nq, self.nr = nq, nr
        self.alndata = np.ctypeslib.as_array(res.alndata, (nq * nr * GTA_ALN_NFIELDS,)).copy() \
            .reshape(nq, nr, GTA_ALN_NFIELDS)
        self.tfms = np.ctypeslib.as_array(res.tfms, (nq * nr * GTA_TFM_NFIELDS,)).copy() \
            .reshape(nq, nr, GTA_TFM_NFIELDS)
        # the alignment section can exceed 2 GiB (C.string_at takes an int size)
        n = int(res.alnsize)
        self.alns = (np.ctypeslib.as_array(C.cast(res.alns, C.POINTER(C.c_uint8)), (n,)).copy() if n
                     else np.zeros(0, np.uint8))
        self.seconds = res.seconds
        self.cells = res.dp_cells


class GtaBatch:
    def __init__(self, device=0, options=None, max_qrys=100, max_qry_poss=50000, max_refs=1 << 20,
                 max_ref_poss=1 << 27):
        self.options = options or default_options()
        h = C.c_void_p()
        rc = lib().gta_create(device, C.byref(self.options), max_qrys, max_qry_poss, max_refs,
                              max_ref_poss, C.byref(h))
        if rc != 0:
            raise GtaError(rc, lib().gta_last_error(None).decode())
        self._h = h
        self._q = self._r = None

    def _check(self, rc):
        if rc != 0:
            raise GtaError(rc, lib().gta_last_error(self._h).decode())

    def TransferQueryPMDataToDevice(self, chunk):
        self._check(lib().gta_set_queries(self._h, C.byref(chunk.view())))
        self._q = chunk

    def TransferCPMDataToDevice(self, chunk):
        self._check(lib().gta_set_references(self._h, C.byref(chunk.view())))
        self._r = chunk

    def ProcessBlock(self):
        res = Results()
        self._check(lib().gta_process(self._h, C.byref(res)))
        return BlockResults(res)

    def process_resident(self):
        """Run the stage leaving results in HBM; returns (device seconds, DP cells)."""
        s, cells = C.c_double(0), C.c_double(0)
        self._check(lib().gta_process_resident(self._h, C.byref(s), C.byref(cells)))
        return s.value, cells.value

    def set_profiling(self, on):
        """Bracket every kernel launch with HIP events (see kernel_stats)."""
        self._check(lib().gta_set_profiling(self._h, 1 if on else 0))

    def kernel_stats(self):
        """Per kernel class of the last step: launches, device seconds, algorithmic bytes."""
        arr = (_KStat * 16)()
        n = lib().gta_kernel_stats(self._h, arr, 16)
        return [{"name": arr[i].name.decode(), "launches": arr[i].launches,
                 "seconds": arr[i].seconds, "alg_bytes": arr[i].alg_bytes, "alg_flops": arr[i].alg_flops}
                for i in range(n)]

    def valu_calibration(self, blocks, iters, kind=0, lanes=64):
        """Duration (s) of the VALU calibration kernel: blocks x 4 waves, each
        issuing 8 * iters VALU instructions of kind `kind` on its first `lanes`
        lanes plus 25 or 26 around the loop (testing hook,
        include/gta/gtalign_amd_testing.h)."""
        s = C.c_double()
        self._check(lib().gta_test_valu_calibration(self._h, blocks, iters, kind, lanes, C.byref(s)))
        return s.value

    def dp_cell_counts(self):
        """DP cells of the last step: (useful Lq x Lr sums, launched lane-steps)."""
        u, l = C.c_double(), C.c_double()
        self._check(lib().gta_dp_cell_counts(self._h, C.byref(u), C.byref(l)))
        return u.value, l.value

    def pair_score(self, d02, d2):
        """The device's pair score d02 / (d02 + d2) for arrays of operands
        (testing hook, include/gta/gtalign_amd_testing.h)."""
        a = np.ascontiguousarray(d02, np.float32)
        b = np.ascontiguousarray(d2, np.float32)
        out = np.empty_like(a)
        self._check(lib().gta_test_pair_score(self._h, a.ctypes.data, b.ctypes.data, a.size, out.ctypes.data))
        return out

    def dp_under(self, tfms, gap0=True, gapopen=0.0, usess=False, dpscan=False):
        """The DP kernel alone under the given per-pair transforms (testing hook,
        include/gta/gtalign_amd_testing.h): returns per pair the matched
        (query, reference) positions in backtrace walk order.  usess=2 runs the
        SS/residue fill of the --add-search-by-ss search instead."""
        nq, nr = self._q.nstrs, self._r.nstrs
        t = np.ascontiguousarray(tfms, np.float32).reshape(nq * nr * GTA_TFM_NFIELDS)
        n = np.zeros(nq * nr, np.int32)
        qp = np.zeros(nq * self._r.nposs, np.int32)
        rp = np.zeros(nq * self._r.nposs, np.int32)
        self._check(lib().gta_test_dp(self._h, t.ctypes.data, int(gap0), gapopen, int(usess),
                                      int(dpscan), n.ctypes.data, qp.ctypes.data, rp.ctypes.data))
        out = {}
        for qi in range(nq):
            for ri in range(nr):
                k = int(n[qi * nr + ri])
                o = qi * self._r.nposs + int(self._r.dst[ri])
                out[qi, ri] = np.stack([qp[o:o + k], rp[o:o + k]], axis=1)
        return out

    def set_dp_buckets(self, min_cells, min_strips):
        """Thresholds of the multi-wave DP bucket (testing hook,
        include/gta/gtalign_amd_testing.h)."""
        self._check(lib().gta_test_set_dp_buckets(self._h, int(min_cells), int(min_strips)))

    def dp_score(self, tfms, slotflags=None, gap0=True, gapopen=0.0):
        """Score-only DP fills (testing hook, include/gta/gtalign_amd_testing.h):
        tfms[k, qi, ri, 12] -> scores[k, qi, ri], the fill's bottom-right cell."""
        nq, nr = self._q.nstrs, self._r.nstrs
        t = np.ascontiguousarray(tfms, np.float32).reshape(-1, nq, nr, GTA_TFM_NFIELDS)
        k = t.shape[0]
        f = None if slotflags is None else np.ascontiguousarray(slotflags, np.int32).reshape(k, nq, nr)
        out = np.zeros((k, nq, nr), np.float32)
        self._check(lib().gta_test_dp_score(self._h, t.ctypes.data, None if f is None else f.ctypes.data, k,
                                            int(gap0), gapopen, out.ctypes.data))
        return out

    def set_autotune(self, on=True, cache_path=None):
        """DP launch-form autotuning per length bucket (gta_set_autotune):
        buckets without a cached choice are timed at the next ProcessBlock;
        choices are read from and appended to cache_path when given."""
        self._check(lib().gta_set_autotune(self._h, int(bool(on)),
                                           None if cache_path is None else str(cache_path).encode()))

    def dp_forms(self):
        """Pairs per DP launch form (one wave, 4, 8, 16 waves, device-wide)
        as of the last call, and the seconds spent autotuning (testing hook)."""
        counts = (C.c_int * 5)()
        secs = C.c_double(0.0)
        self._check(lib().gta_test_dp_forms(self._h, counts, C.byref(secs)))
        return {"one_wave": counts[0], "w4": counts[1], "w8": counts[2], "w16": counts[3],
                "grid": counts[4]}, secs.value

    def set_dp_grid(self, mode):
        """Device-wide strip pipeline for the multi-wave DP bucket: 0 when
        those pairs number at most half the CUs, 1 always, 2 never (testing
        hook, include/gta/gtalign_amd_testing.h)."""
        self._check(lib().gta_test_set_dp_grid(self._h, int(mode)))

    def set_split(self, min_len, seeds_per_part=256, slots_per_part=32):
        """Workgroup split of the superposition searches for long pairs
        (testing hook, include/gta/gtalign_amd_testing.h)."""
        self._check(lib().gta_test_set_split(self._h, int(min_len), int(seeds_per_part),
                                             int(slots_per_part)))

    def close(self):
        if getattr(self, "_h", None):
            lib().gta_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

