"""FP64 restatements used as checks (TEST INFRASTRUCTURE ONLY): rigid motions,
SVD superposition, TM-score from an alignment and a transform, and readers of
the alignment section.  Used by tests/ and bench.py's CPU-baseline leg."""
import numpy as np

from gtalign_alpha_amd.capi import ALN_BEGCOORDS, ALN_ALNLENGTH


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def kabsch_svd(x, y):
    """Least-squares rotation/translation x -> y in FP64 (SVD)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    xc, yc = x.mean(0), y.mean(0)
    h = (x - xc).T @ (y - yc)
    u, s, vt = np.linalg.svd(h)
    d = np.sign(np.linalg.det(vt.T @ u.T))
    D = np.diag([1.0, 1.0, d])
    R = vt.T @ D @ u.T
    return R, yc - R @ xc


def tfm_apply(t, x):
    R = np.asarray(t[:9], np.float64).reshape(3, 3)
    return np.asarray(x, np.float64) @ R.T + np.asarray(t[9:12], np.float64)


def d0_tm(L):
    """TM-score d0 (Zhang & Skolnick 2004) with the 0.5 floor used for output."""
    return max(0.5, 1.24 * np.cbrt(L - 15.0) - 1.8) if L > 21 else 0.5


def alignment_rows(alns, q, r, qi, ri):
    """The 5 alignment rows of pair (qi, ri) from the alignment section."""
    nr = r.nstrs
    qlen, qd = int(q.len[qi]), int(q.dst[qi])
    off = (qi * r.nposs + nr * (qd + qi)) * 5
    rowlen = r.nposs + nr * (qlen + 1)
    beg = int(r.dst[ri]) + ri * (qlen + 1)
    seg = qlen + int(r.len[ri]) + 1
    rows = []
    for f in range(5):
        b = bytes(alns[off + beg + rowlen * f: off + beg + rowlen * f + seg])
        rows.append(b.split(b"\0")[0].decode())
    return rows


def matched_positions(rows, alndata_pair):
    """Matched (query, reference) positions of an alignment from its rows and
    1-based begin coordinates."""
    beg = np.float32(alndata_pair[ALN_BEGCOORDS]).view(np.uint32)
    qp, rp = int(beg & 0xffff) - 1, int(beg >> 16) - 1
    out = []
    for a, b in zip(rows[0], rows[2]):
        if a != '-' and b != '-':
            out.append((qp, rp))
        if a != '-':
            qp += 1
        if b != '-':
            rp += 1
    assert len(rows[0]) == int(alndata_pair[ALN_ALNLENGTH])
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def tm_fp64(q, r, qi, ri, pairs, tfm, L):
    """TM-score in FP64 of matched pairs under tfm, normalised by L with d0(L)."""
    qx = q.coords(qi)[pairs[:, 0]]
    rx = r.coords(ri)[pairs[:, 1]]
    d2 = ((tfm_apply(tfm, qx) - rx) ** 2).sum(1)
    d0 = d0_tm(L)
    return float((d0 * d0 / (d0 * d0 + d2)).sum() / L)


def fp64_errors(res, q, r):
    """Per aligned pair of a result: |reported TM - FP64 TM| (both
    normalisations, under the reported transform) and |reported RMSD - FP64
    RMSD| (SVD superposition of the aligned pairs).  Returns two arrays."""
    from gtalign_alpha_amd.capi import ALN_RMSD, ALN_SCOREQ, ALN_SCORER
    dtm, drm = [], []
    for qi in range(q.nstrs):
        for ri in range(r.nstrs):
            a = res.alndata[qi, ri]
            if a[ALN_RMSD] >= 9999:
                continue
            pairs = matched_positions(alignment_rows(res.alns, q, r, qi, ri), a)
            tq = tm_fp64(q, r, qi, ri, pairs, res.tfms[qi, ri], int(q.len[qi]))
            tr = tm_fp64(q, r, qi, ri, pairs, res.tfms[qi, ri], int(r.len[ri]))
            dtm.append(max(abs(tq - a[ALN_SCOREQ]), abs(tr - a[ALN_SCORER])))
            x, y = q.coords(qi)[pairs[:, 0]], r.coords(ri)[pairs[:, 1]]
            R, t = kabsch_svd(x, y)
            rms = float(np.sqrt(((x @ R.T + t - y) ** 2).sum(1).mean()))
            drm.append(abs(rms - a[ALN_RMSD]))
    return np.array(dtm), np.array(drm)
