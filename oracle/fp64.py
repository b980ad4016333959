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


def matched_positions_np(alns, alndata, q, r, qi, ri):
    """matched_positions(alignment_rows(...)) with numpy on the row bytes."""
    nr = r.nstrs
    qlen, qd = int(q.len[qi]), int(q.dst[qi])
    off = (qi * r.nposs + nr * (qd + qi)) * 5
    rowlen = r.nposs + nr * (qlen + 1)
    beg = int(r.dst[ri]) + ri * (qlen + 1)
    n = int(alndata[ALN_ALNLENGTH])
    a = alns[off + beg: off + beg + n]
    b = alns[off + beg + 2 * rowlen: off + beg + 2 * rowlen + n]
    gap = ord('-')
    qa, ra = a != gap, b != gap
    bc = np.float32(alndata[ALN_BEGCOORDS]).view(np.uint32)
    qp = np.cumsum(qa) - 1 + (int(bc & 0xffff) - 1)
    rp = np.cumsum(ra) - 1 + (int(bc >> 16) - 1)
    m = qa & ra
    return np.stack([qp[m], rp[m]], 1).astype(np.int64)


def fp64_errors_all(res, q, r):
    """fp64_errors over every aligned pair of a large result, with the rows
    parsed by numpy: per aligned pair |reported TM - FP64 TM| (both
    normalisations, reported transform) and |reported RMSD - FP64 RMSD|."""
    from gtalign_alpha_amd.capi import ALN_RMSD, ALN_SCOREQ, ALN_SCORER
    alns = np.frombuffer(res.alns, np.uint8) if not isinstance(res.alns, np.ndarray) else res.alns.view(np.uint8)
    dtm, drm = [], []
    for qi, ri in np.argwhere(res.alndata[..., ALN_RMSD] < 9999):
        a = res.alndata[qi, ri]
        pairs = matched_positions_np(alns, a, q, r, qi, ri)
        x, y = q.coords(qi)[pairs[:, 0]].astype(np.float64), r.coords(ri)[pairs[:, 1]].astype(np.float64)
        d2 = ((tfm_apply(res.tfms[qi, ri], x) - y) ** 2).sum(1)
        e = 0.0
        for L, f in ((int(q.len[qi]), ALN_SCOREQ), (int(r.len[ri]), ALN_SCORER)):
            d0 = d0_tm(L)
            e = max(e, abs(float((d0 * d0 / (d0 * d0 + d2)).sum() / L) - a[f]))
        dtm.append(e)
        R, t = kabsch_svd(x, y)
        drm.append(abs(float(np.sqrt(((x @ R.T + t - y) ** 2).sum(1).mean())) - a[ALN_RMSD]))
    return np.array(dtm), np.array(drm)
