"""Seeded synthetic workloads (the benchmark configurations of the stage).

Chains come from gta_synth_chain (csrc/synth.cpp): 5-20 residue helix,
strand and coil segments inside a sphere of radius 2.2*L^0.38 A.  Homologs
come from gta_synth_homolog: a source chain with deletions, short coil
insertions, Gaussian noise, a random rotation and a translation.
"""
import ctypes as C
import os

import numpy as np

from .chunk import Chunk

_HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_LIB = os.path.join(_HERE, "libgta_synth.so")
AA = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
BASE_SEED = 20261202

_lib = None


def _synth():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_LIB):
            raise RuntimeError("%s is not built; run __graft_entry__.build()" % SYNTH_LIB)
        _lib = C.CDLL(SYNTH_LIB)
        _lib.gta_synth_chain.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        _lib.gta_synth_chain.restype = None
        _lib.gta_synth_homolog.argtypes = [C.c_uint64, C.c_void_p, C.c_int, C.c_float, C.c_float,
                                           C.c_float, C.c_float, C.c_void_p, C.c_int, C.c_void_p]
        _lib.gta_synth_homolog.restype = C.c_int
        _lib.gta_synth_chain_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                               C.c_int]
        _lib.gta_synth_chain_batch.restype = None
    return _lib


def chains(seeds, lens):
    """chain(seeds[k], lens[k]) for every k, generated on the host's threads."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.int32)
    offs = np.zeros(len(lens), dtype=np.int64)
    if len(lens) > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    out = np.empty((int(lens.sum()), 3), dtype=np.float32)
    nthreads = max(1, min(16, os.cpu_count() or 1))
    _synth().gta_synth_chain_batch(seeds.ctypes.data, lens.ctypes.data, offs.ctypes.data, len(lens),
                                   out.ctypes.data, nthreads)
    return [out[o:o + L] for o, L in zip(offs, lens)]


def chain(seed, L):
    out = np.empty((L, 3), dtype=np.float32)
    _synth().gta_synth_chain(C.c_uint64(seed), int(L), out.ctypes.data)
    return out


def homolog(seed, src, sigma, maxtrans, delfrac, insfrac):
    src = np.ascontiguousarray(src, dtype=np.float32)
    cap = 2 * len(src) + 8
    out = np.empty((cap, 3), dtype=np.float32)
    idx = np.empty(cap, dtype=np.int32)
    n = _synth().gta_synth_homolog(C.c_uint64(seed), src.ctypes.data, len(src), float(sigma),
                                   float(maxtrans), float(delfrac), float(insfrac), out.ctypes.data,
                                   cap, idx.ctypes.data)
    return out[:n].copy(), idx[:n].copy()


def lognormal_lengths(rng, n, median, sigma, lo, hi):
    return np.clip(np.round(np.exp(rng.normal(np.log(median), sigma, n))), lo, hi).astype(np.int64)


def residues(rng, L):
    return AA[rng.integers(0, len(AA), L)].tobytes().decode()


def _hom_residues(rng, src_res, idx):
    out = np.frombuffer(residues(rng, len(idx)).encode(), dtype=np.uint8).copy()
    src = np.frombuffer(src_res.encode(), dtype=np.uint8)
    keep = idx >= 0
    out[keep] = src[idx[keep]]
    return out.tobytes().decode()


def make_set(seed, n, lens, queries, qres, hom_frac, sigma_lo, sigma_hi, del_hi, ins_hi,
             maxtrans=50.0):
    """n reference chains; a fraction hom_frac are homologs of a random query."""
    rng = np.random.default_rng(seed)
    structs, res, src = [], [], np.full(n, -1, dtype=np.int64)
    later, lseeds, llens = [], [], []     # unrelated chains, generated in one batch
    for i in range(n):
        s = int(rng.integers(0, 2**62))
        if len(queries) and rng.random() < hom_frac:
            qi = int(rng.integers(0, len(queries)))
            sig = rng.uniform(sigma_lo, sigma_hi)
            xyz, idx = homolog(s, queries[qi], sig, rng.uniform(0, maxtrans),
                               rng.uniform(0, del_hi), rng.uniform(0, ins_hi))
            structs.append(xyz)
            res.append(_hom_residues(rng, qres[qi], idx))
            src[i] = qi
        else:
            L = int(lens[i])
            structs.append(None)
            later.append(i)
            lseeds.append(s)
            llens.append(L)
            res.append(residues(rng, L))
    for i, xyz in zip(later, chains(np.array(lseeds, dtype=np.uint64), llens)):
        structs[i] = xyz
    return structs, res, src


def stress_chain(rng, L, straight_frac=0.4):
    """A chain of 3-8 residue segments, a fraction straight_frac of them
    straight (collinear CA atoms 3.8 A apart) and the rest ideal helix (rise
    1.5 A, 100 deg per residue, radius 2.3 A), joined at random angles: most
    superposition fragments of such a chain are near-collinear or near-planar."""
    out = np.zeros((L, 3))
    pos = np.zeros(3)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    i = 0
    while i < L:
        n = min(int(rng.integers(3, 9)), L - i)
        # a new direction at a random angle (30-110 deg) from the previous one
        v = rng.normal(size=3)
        v -= v.dot(d) * d
        v /= np.linalg.norm(v)
        ang = np.deg2rad(rng.uniform(30, 110))
        d = np.cos(ang) * d + np.sin(ang) * v
        d /= np.linalg.norm(d)
        if rng.random() < straight_frac:
            for k in range(n):
                pos = pos + 3.8 * d
                out[i + k] = pos
        else:
            u = np.cross(d, v)
            u /= np.linalg.norm(u)
            c0 = pos - 2.3 * v
            for k in range(n):
                t = np.deg2rad(100.0) * (k + 1)
                out[i + k] = c0 + 2.3 * (np.cos(t) * v + np.sin(t) * u) + 1.5 * (k + 1) * d
            pos = out[i + n - 1]
        i += n
    return (out - out.mean(0)).astype(np.float32)


def rigid_copy(rng, x, sigma, maxtrans=50.0):
    """x under a random rotation and translation with Gaussian noise sigma."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, e = q
    R = np.array([[a * a + b * b - c * c - e * e, 2 * (b * c - a * e), 2 * (b * e + a * c)],
                  [2 * (b * c + a * e), a * a - b * b + c * c - e * e, 2 * (c * e - a * b)],
                  [2 * (b * e - a * c), 2 * (c * e + a * b), a * a - b * b - c * c + e * e]])
    t = rng.uniform(-maxtrans, maxtrans, 3)
    y = np.asarray(x, np.float64) @ R.T + t + rng.normal(scale=sigma, size=x.shape)
    return y.astype(np.float32)


def assembly(rng, ndom=150, total=37860, min_sep=40.0):
    """ndom generated domains (lengths summing to total, ~total/ndom each)
    placed by random rigid motions with domain centres at least min_sep A
    apart (no overlap), concatenated into one structure."""
    base, extra = divmod(total, ndom)
    lens = [base + 1] * extra + [base] * (ndom - extra)
    side = (ndom * min_sep ** 3) ** (1.0 / 3.0) * 1.6
    centres = []
    while len(centres) < ndom:
        c = rng.uniform(0, side, 3)
        if all(np.linalg.norm(c - d) >= min_sep for d in centres):
            centres.append(c)
    parts = []
    for L, c in zip(lens, centres):
        x = chain(int(rng.integers(0, 2**62)), L).astype(np.float64)
        x -= x.mean(0)
        parts.append(rigid_copy(rng, x, 0.0, maxtrans=0.0).astype(np.float64) + c)
    return np.concatenate(parts).astype(np.float32)


def config(index, scale=1.0, seed=None):
    """Query and reference chunks of one benchmark configuration.

    index 0: 1 query of 150 vs 256 refs U(50,300), 30% homologs (sigma 1, 0-10% deletions)
    index 1: 16 queries x 50,000 refs, lognormal(250, 0.5) in [30,1000], 20% homologs
    index 2: 1 query of 400 x 500,000 refs, lognormal(220, 0.7) in [20,4000], 10% homologs
    index 3: precision stress, 20 query chains of straight and helical 3-8 residue
             segments (lengths 60-300) x 2,000 references; reference j is a copy of
             query j mod 20 under a random rigid motion with noise sigma 0.01-0.1 A
    index 4: all-vs-all family set: 500 families x 40 chains (seed lengths lognormal(200, 0.5)
             in [40, 800]; members under rigid motions, noise 0.5-3 A, 0-15% indels) = 20,000
             references; 16 queries (seed chains of evenly spaced families; scale < 1 gives
             fewer queries)
    `scale` shrinks the number of references (tests and bounded CPU samples).
    Returns (queries, references, src): src[j] is the query-chunk position of
    the query that reference-chunk position j is a homolog (copy) of, or -1.
    """
    seed = BASE_SEED + index if seed is None else seed
    rng = np.random.default_rng(seed)
    if index == 0:
        nq, nr = 1, max(1, int(round(256 * scale)))
        qlens = np.array([150])
        rlens = rng.integers(50, 301, nr)
        hom, slo, shi, dhi, ihi = 0.3, 1.0, 1.0, 0.10, 0.0
    elif index == 1:
        nq, nr = 16, max(1, int(round(50000 * scale)))
        qlens = lognormal_lengths(rng, nq, 250, 0.5, 30, 1000)
        rlens = lognormal_lengths(rng, nr, 250, 0.5, 30, 1000)
        hom, slo, shi, dhi, ihi = 0.2, 0.5, 3.0, 0.075, 0.075
    elif index == 2:
        nq, nr = 1, max(1, int(round(500000 * scale)))
        qlens = np.array([400])
        rlens = lognormal_lengths(rng, nr, 220, 0.7, 20, 4000)
        hom, slo, shi, dhi, ihi = 0.1, 0.5, 3.0, 0.075, 0.075
    elif index == 4:
        # 500 families of 1 seed chain + 39 members (rigid motion, noise 0.5-3 A,
        # 0-7.5% deletions and 0-7.5% insertions); all 20,000 chains are the
        # references and the first chain of every 500/nq-th family a query
        nfam, per = 500, 40
        nq = max(1, int(round(16 * scale))) if scale < 1.0 else 16
        seeds = lognormal_lengths(rng, nfam, 200, 0.5, 40, 800)
        fs = chains(rng.integers(0, 2**62, nfam).astype(np.uint64), seeds)
        fr = [residues(rng, int(L)) for L in seeds]
        rs, rr, src = [], [], []
        for f in range(nfam):
            rs.append(fs[f]); rr.append(fr[f]); src.append(f)
            for m in range(per - 1):
                xyz, idx = homolog(int(rng.integers(0, 2**62)), fs[f], rng.uniform(0.5, 3.0), 50.0,
                                   rng.uniform(0, 0.075), rng.uniform(0, 0.075))
                rs.append(xyz); rr.append(_hom_residues(rng, fr[f], idx)); src.append(f)
        qf = np.linspace(0, nfam - 1, nq).astype(int)
        qs = [fs[f] for f in qf]
        qr = [fr[f] for f in qf]
        fam_to_q = np.full(nfam, -1, dtype=np.int64)
        fam_to_q[qf] = np.arange(nq)
        return _chunks(qs, qr, rs, rr, fam_to_q[np.asarray(src)])
    elif index == 3:
        nq, nr = 20, max(1, int(round(2000 * scale)))
        qs = [stress_chain(rng, int(L)) for L in rng.integers(60, 301, nq)]
        qr = [residues(rng, len(x)) for x in qs]
        rs, rr, src = [], [], np.zeros(nr, dtype=np.int64)
        for j in range(nr):
            src[j] = j % nq
            rs.append(rigid_copy(rng, qs[src[j]], rng.uniform(0.01, 0.1)))
            rr.append(qr[src[j]])
        return _chunks(qs, qr, rs, rr, src)
    else:
        raise ValueError("unsupported configuration %d" % index)
    qs = [chain(int(rng.integers(0, 2**62)), int(L)) for L in qlens]
    qr = [residues(rng, int(L)) for L in qlens]
    rs, rr, src = make_set(int(rng.integers(0, 2**62)), nr, rlens, qs, qr, hom, slo, shi, dhi, ihi)
    return _chunks(qs, qr, rs, rr, src)


def _chunks(qs, qr, rs, rr, src):
    """Chunks sorted by length, with src mapped to chunk positions."""
    qc = Chunk.from_structures(qs, qr)
    rc = Chunk.from_structures(rs, rr)
    qpos = np.full(len(qs) + 1, -1, dtype=np.int64)
    qpos[qc.order] = np.arange(len(qs))
    s = np.asarray(src)[rc.order]
    return qc, rc, np.where(s >= 0, qpos[s], -1)


# ---- nucleic acids ---------------------------------------------------------------
# C3' traces: consecutive C3' atoms ~5.9 A apart; a base pair puts its two C3'
# atoms ~13.75 A apart (the reference's C3' bounds 12.5-15 A,
# PM2DVectorFields.h:178-206).
NA_BASES = np.frombuffer(b"ACGU", dtype=np.uint8)
_COMPLEMENT = {ord("A"): ord("U"), ord("U"): ord("A"), ord("G"): ord("C"), ord("C"): ord("G")}


def na_compact(seed, L):
    """A compact C3' trace of L residues (a generated chain scaled to 5.9 A steps)."""
    return chain(seed, L) * np.float32(5.9 / 3.8)


def na_hairpin(rng, nstem, nloop, noise=0.3):
    """A stem-loop C3' trace: nstem residues up one strand, a loop of nloop
    residues, nstem residues back down the other strand 13.75 A across; the
    second strand's bases complement the first's.  Returns (coords, bases)."""
    up = np.stack([np.zeros(nstem), 5.9 * np.arange(nstem), np.zeros(nstem)], 1)
    down = up[::-1] + np.array([13.75, 0.0, 0.0])
    ang = np.linspace(0, np.pi, nloop + 2)[1:-1]
    loop = np.stack([6.875 - 6.875 * np.cos(ang), 5.9 * (nstem - 1) + 6.0 * np.sin(ang), np.zeros(nloop)], 1)
    xyz = np.concatenate([up, loop, down]) + rng.normal(scale=noise, size=(2 * nstem + nloop, 3))
    b1 = NA_BASES[rng.integers(0, 4, nstem)]
    bl = NA_BASES[rng.integers(0, 4, nloop)]
    b2 = np.array([_COMPLEMENT[int(b)] for b in b1[::-1]], dtype=np.uint8)
    return xyz.astype(np.float32), np.concatenate([b1, bl, b2]).tobytes().decode()


def na_bases(rng, L):
    return NA_BASES[rng.integers(0, 4, L)].tobytes().decode()


def mark_nucleic(c, nucleic):
    """Sets the insertion/chain/order words of chunk c: positive (nucleic acid,
    PM2DVectorFields.h:96-99) for the structures at the chunk positions where
    nucleic[i] is true, negative (protein) for the others; chain 'A', order 1."""
    word = (ord(" ") << 16) | (ord("A") << 8) | 1
    for i in range(c.nstrs):
        d, l = int(c.dst[i]), int(c.len[i])
        c.ins_ch_ord[d:d + l] = word if nucleic[i] else -word
    c._view = None
    return c
