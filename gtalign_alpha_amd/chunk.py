"""Structure chunks in the reference's SoA chunk format.

A chunk holds the 10 field arrays of TPM2DVectorFields
(src/libgenp/gdats/PM2DVectorFields.h:220-258): per structure its length,
prefix offset and type; per residue x, y, z (one array per coordinate),
the insertion/chain/order word, the residue number, the residue letter and
the secondary-structure letter.  Structures are kept sorted by descending
length, as PMBatchStrData::Sort does (PMBatchStrData.cpp:110-160); `order`
maps each chunk position back to the caller's structure index.
"""
import ctypes as C

import numpy as np

from .capi import ChunkView

MAX_STRUCTURE_LENGTH = 65535   # PMBatchStrData.h:34 (16-bit DP cell coordinates)


class Chunk:
    def __init__(self, lens, x, y, z, rsd=None, ins_ch_ord=None, order=None):
        self.len = np.ascontiguousarray(lens, dtype=np.int32)
        n = int(self.len.sum())
        self.dst = np.zeros(len(self.len), dtype=np.uint32)
        if len(self.len) > 1:
            self.dst[1:] = np.cumsum(self.len[:-1], dtype=np.uint64).astype(np.uint32)
        self.type = np.zeros(len(self.len), dtype=np.int32)
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.y = np.ascontiguousarray(y, dtype=np.float32)
        self.z = np.ascontiguousarray(z, dtype=np.float32)
        assert self.x.shape == (n,) and self.y.shape == (n,) and self.z.shape == (n,)
        self.ins_ch_ord = (np.zeros(n, dtype=np.int32) if ins_ch_ord is None
                           else np.ascontiguousarray(ins_ch_ord, dtype=np.int32))
        self.resnum = np.zeros(n, dtype=np.int32)
        if len(self.len):
            self.resnum = (np.arange(n, dtype=np.int64) -
                           np.repeat(self.dst.astype(np.int64), self.len)).astype(np.int32) + 1
        self.rsd = (np.full(n, ord('A'), dtype=np.uint8) if rsd is None
                    else np.ascontiguousarray(rsd, dtype=np.uint8))
        self.ss = np.full(n, ord(' '), dtype=np.uint8)
        self.order = (np.arange(len(self.len)) if order is None else np.asarray(order))
        self._view = None

    @property
    def nstrs(self):
        return int(len(self.len))

    @property
    def nposs(self):
        return int(self.len.sum())

    def coords(self, i):
        d, l = int(self.dst[i]), int(self.len[i])
        return np.stack([self.x[d:d + l], self.y[d:d + l], self.z[d:d + l]], axis=1)

    def residues(self, i):
        d, l = int(self.dst[i]), int(self.len[i])
        return self.rsd[d:d + l].tobytes().decode()

    def view(self):
        """gta_chunk_view over this chunk's arrays (valid while self lives)."""
        if self._view is None:
            P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
            self._view = ChunkView(
                nstrs=self.nstrs, nposs=self.nposs,
                len=P(self.len, C.c_int), dst=P(self.dst, C.c_uint), type=P(self.type, C.c_int),
                x=P(self.x, C.c_float), y=P(self.y, C.c_float), z=P(self.z, C.c_float),
                ins_ch_ord=P(self.ins_ch_ord, C.c_int), resnum=P(self.resnum, C.c_int),
                rsd=C.cast(self.rsd.ctypes.data, C.c_char_p), ss=None)
        return self._view

    @staticmethod
    def from_structures(structs, residues=None, sort=True):
        """Chunk from a list of (L,3) coordinate arrays, sorted by descending length
        (stable for equal lengths)."""
        lens = np.array([len(s) for s in structs], dtype=np.int64)
        if np.any(lens < 1):
            raise ValueError("empty structure")
        if np.any(lens > MAX_STRUCTURE_LENGTH):
            raise ValueError("structure longer than %d residues" % MAX_STRUCTURE_LENGTH)
        order = np.argsort(-lens, kind="stable") if sort else np.arange(len(structs))
        xyz = np.concatenate([np.asarray(structs[i], dtype=np.float32) for i in order], axis=0) \
            if len(structs) else np.zeros((0, 3), np.float32)
        rsd = None
        if residues is not None:
            rsd = np.frombuffer("".join(residues[i] for i in order).encode(), dtype=np.uint8)
        return Chunk(lens[order], xyz[:, 0], xyz[:, 1], xyz[:, 2], rsd=rsd, order=order)

    def subset(self, idx):
        """Chunk of the structures at chunk positions idx (kept in that order)."""
        structs = [self.coords(i) for i in idx]
        res = [self.residues(i) for i in idx]
        c = Chunk.from_structures(structs, res, sort=False)
        c.order = self.order[np.asarray(idx)]
        return c
