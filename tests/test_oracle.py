"""Pins of the oracle (CPU): closed forms, brute force, and the reference's own
CPU code where it was built (oracle/_ref)."""
import itertools

import numpy as np
import pytest

from gtalign_alpha_amd import synth
from gtalign_alpha_amd.capi import default_options
from oracle import oracle
from oracle.fp64 import random_rotation, kabsch_svd, tfm_apply, d0_tm


# ---- Kabsch (transformbase.h:167-583) ---------------------------------------------

def test_kabsch_recovers_rigid_motion():
    rng = np.random.default_rng(1)
    for n in (3, 4, 10, 200):
        x = rng.normal(scale=10.0, size=(n, 3))
        R = random_rotation(rng)
        t = rng.normal(scale=30.0, size=3)
        y = x @ R.T + t
        tf = oracle.kabsch(x, y)
        Rk = tf[:9].reshape(3, 3).astype(np.float64)
        assert np.allclose(Rk, R, atol=2e-3), n
        assert np.allclose(tf[9:], t, atol=2e-2), n
        assert np.allclose(Rk @ Rk.T, np.eye(3), atol=1e-4)
        assert abs(np.linalg.det(Rk) - 1.0) < 1e-4


def test_kabsch_matches_svd_on_noisy_data():
    rng = np.random.default_rng(2)
    for _ in range(20):
        n = int(rng.integers(5, 300))
        x = rng.normal(scale=12.0, size=(n, 3))
        y = x @ random_rotation(rng).T + rng.normal(scale=20, size=3) + rng.normal(scale=2.0, size=(n, 3))
        tf = oracle.kabsch(x, y)
        R, t = kabsch_svd(x, y)
        # compare the residuals: both are least-squares optima
        e_orc = np.sqrt((((tfm_apply(tf, x) - y) ** 2).sum(1)).mean())
        e_svd = np.sqrt((((x @ R.T + t - y) ** 2).sum(1)).mean())
        assert e_orc <= e_svd * (1 + 1e-3) + 1e-4


def test_kabsch_rejects_reflection():
    # a mirrored copy must still give a proper rotation (det = +1)
    rng = np.random.default_rng(3)
    x = rng.normal(scale=10.0, size=(50, 3))
    y = x * np.array([1.0, 1.0, -1.0])
    Rk = oracle.kabsch(x, y)[:9].reshape(3, 3).astype(np.float64)
    assert abs(np.linalg.det(Rk) - 1.0) < 1e-4


def test_kabsch_degenerate_inputs_finite():
    # collinear and coincident points: the reference falls back to a valid transform
    x = np.stack([np.arange(6.0), np.zeros(6), np.zeros(6)], 1)
    for y in (x + 3.0, np.zeros((6, 3)), x[::-1].copy()):
        tf = oracle.kabsch(x, y)
        assert np.all(np.isfinite(tf))


def test_rmsd_e0_identity_matches_direct():
    rng = np.random.default_rng(4)
    for _ in range(10):
        n = int(rng.integers(10, 300))
        x = rng.normal(scale=10.0, size=(n, 3))
        y = x @ random_rotation(rng).T + rng.normal(scale=1.5, size=(n, 3))
        R, t = kabsch_svd(x, y)
        direct = np.sqrt(((x @ R.T + t - y) ** 2).sum(1).mean())
        # E0 identity in FP32 loses precision by cancellation; 1e-2 A bounds it here
        assert abs(oracle.rmsd(x, y) - direct) < 1e-2


# ---- d0 / fragment geometry (stagecnsts.cuh:48-175, fragment.cuh:52-71) ------------

def test_d0_textbook_values():
    for L in (22, 50, 100, 300, 1000):
        d = oracle.d0(L, L + 7)
        ref = 1.24 * np.cbrt(L - 15.0) - 1.8
        assert abs(d["d0fin"] - ref) < 1e-5
        assert abs(d["d0"] - (ref + 0.8)) < 1e-5
        assert abs(d["d02_dpscan"] - (ref + 2.3) ** 2) < 1e-4
        assert abs(d["d82"] - (1.5 * L ** 0.3 + 3.5) ** 2) < 1e-3
        assert 4.5 <= d["d0s"] <= 8.0
    assert abs(oracle.d0(10, 10)["d0fin"] - 0.5) < 1e-7
    assert abs(oracle.d0(10, 10)["d0"] - (0.168 + 0.8)) < 1e-6


def test_fraglen_halving_and_floor():
    # fragment lengths halve with the index and never go below 4 (or 0 when too short)
    for L in (5, 8, 9, 40, 100, 257):
        lens = [oracle.fraglen(L, L, 0, 0, k) for k in range(6)]
        assert lens[0] == L
        for k in range(1, 6):
            assert lens[k] == 0 or lens[k] >= 4
            assert lens[k] <= max(4, L >> k) or lens[k] == 0
        assert lens[5] in (0, 4)
    assert oracle.fraglen(3, 3, 0, 0, 0) == 3
    assert oracle.fraglen(50, 80, 10, 0, 0) == 40


# ---- DP (MpDPHub.cpp:73-388) ------------------------------------------------------

def _paths(Lq, Lr):
    """All monotone paths from the boundary to (Lq-1, Lr-1) as lists of moves
    into cells; a path may start at any boundary point (free leading gaps)."""
    out = []

    def rec(i, j, acc):
        if i < 0 or j < 0:
            out.append(list(reversed(acc)))
            return
        rec(i - 1, j - 1, acc + [("D", i, j)])
        rec(i, j - 1, acc + [("L", i, j)])
        rec(i - 1, j, acc + [("U", i, j)])

    rec(Lq - 1, Lr - 1, [])
    return out


def test_dp_gap0_is_the_optimum_over_paths():
    """Without gap costs the fill is a plain max over monotone paths; brute force
    over every path of small matrices must reach the same optimum."""
    rng = np.random.default_rng(5)
    I = np.eye(3).reshape(-1).tolist() + [0, 0, 0]
    for Lq, Lr in [(3, 4), (4, 4), (5, 3), (4, 6)]:
        for _ in range(3):
            q = rng.normal(scale=4.0, size=(Lq, 3))
            r = rng.normal(scale=4.0, size=(Lr, 3))
            d02 = 4.0
            _, pairs = oracle.dp(q, r, I, d02, 0.0, True)
            d2 = ((q[:, None, :].astype(np.float32) - r[None, :, :].astype(np.float32)) ** 2).sum(-1)
            s = d02 / (d02 + d2)
            best = max(sum(s[i, j] for m, i, j in p if m == "D") for p in _paths(Lq, Lr))
            got = sum(s[i, j] for i, j in pairs)
            assert abs(best - got) < 1e-5, (Lq, Lr)


def test_dp_score_is_the_optimum_over_paths():
    """The score-only fill (GAP0, the only form the search uses:
    ExecDPScore128xKernel<true>, MpStageFrg3.h:83-87) returns the bottom-right
    cell: the maximum over every monotone path of the summed match scores
    d0^2/(d0^2+d^2) with the search d0 of min(Lq, Lr), reached by brute force on
    3-6 residue matrices; without a gap cost no cell is stored negated, so the
    value is non-negative, and it grows when a residue pair is moved closer."""
    rng = np.random.default_rng(17)
    for Lq, Lr in [(3, 3), (3, 5), (4, 4), (5, 3), (4, 6), (6, 4)]:
        for _ in range(3):
            q = rng.normal(scale=4.0, size=(Lq, 3)).astype(np.float32)
            r = rng.normal(scale=4.0, size=(Lr, 3)).astype(np.float32)
            R = random_rotation(rng)
            t = rng.normal(scale=1.0, size=3)
            tfm = np.concatenate([R.reshape(-1), t]).astype(np.float32)
            d0 = np.float32(oracle.d0(Lq, Lr)["d0"])
            d02 = np.float64(d0) * np.float64(d0)
            qt = q.astype(np.float64) @ R.T.astype(np.float32).astype(np.float64) + tfm[9:].astype(np.float64)
            d2 = ((qt[:, None, :] - r[None, :, :].astype(np.float64)) ** 2).sum(-1)
            s = d02 / (d02 + d2)
            best = max(sum(s[i, j] for m, i, j in p if m == "D") for p in _paths(Lq, Lr))
            got = oracle.dp_score(q, r, tfm)
            assert got >= 0.0
            assert abs(best - got) < 1e-4 * max(1.0, best), (Lq, Lr, best, got)
    # a lower bound that any wrong cell would break: the diagonal path alone
    c = synth.chain(31, 6)
    I = np.eye(3).reshape(-1).tolist() + [0, 0, 0]
    assert abs(oracle.dp_score(c, c, I) - 6.0) < 1e-5    # identical chains: every pair scores 1


def test_dp_identical_chains_align_diagonally():
    c = synth.chain(11, 80)
    I = np.eye(3).reshape(-1).tolist() + [0, 0, 0]
    for gap0, g in ((True, 0.0), (False, -0.6)):
        _, pairs = oracle.dp(c, c, I, 9.0, g, gap0)
        assert len(pairs) == 80
        assert np.all(pairs[:, 0] == pairs[:, 1])


@pytest.mark.parametrize("gap0,gapopen", [(True, 0.0), (False, -0.6), (False, -1.0)])
def test_dp_matches_reference_cell_by_cell(have_ref, gap0, gapopen):
    """The oracle's backtrace matrix equals the reference's MpDPHub matrix,
    for random superpositions of synthetic chains."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    rng = np.random.default_rng(6)
    for k in range(6):
        Lq, Lr = int(rng.integers(20, 140)), int(rng.integers(20, 140))
        q = synth.chain(100 + k, Lq)
        r = synth.chain(200 + k, Lr)
        R = random_rotation(rng)
        t = rng.normal(scale=3.0, size=3)
        tfm = np.concatenate([R.reshape(-1), t]).astype(np.float32)
        d = oracle.d0(Lq, Lr)
        bo, po = oracle.dp(q, r, tfm, float(d["d0"]) ** 2, gapopen, gap0)
        br, pr = oracle.ref_dp(q, r, tfm, gap0, gapopen)
        assert np.array_equal(bo, br), (k, Lq, Lr)
        assert np.array_equal(po, pr), (k, Lq, Lr)


# ---- the search by secondary structure and residue similarity -------------------------

def test_residue_scores_equal_reference_table(have_ref):
    """The Gonnet table in bits (the HIP library uploads the same table)
    equals the reference's GONNET_SCORES entry for entry."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    assert np.array_equal(oracle.gonnet(), oracle.ref_gonnet())


def test_residue_scores_known_values():
    """Entries fixed by the published matrix: log2(10^(G/10)) for G the
    Gonnet entry (A-A 2.4, W-W 14.2, C-C 11.5, W-G -4.0), symmetric, and 0
    for letters outside the 20 amino acids."""
    g = oracle.gonnet()
    idx = lambda a: ord(a) - ord("A")
    for a, b, v in [("A", "A", 2.4), ("W", "W", 14.2), ("C", "C", 11.5), ("W", "G", -4.0), ("L", "I", 2.8)]:
        assert abs(g[idx(a), idx(b)] - v / 10.0 * np.log2(10.0)) < 1e-5, (a, b)
    assert np.array_equal(g, g.T)
    for x in "BJOUXZ":
        assert not g[idx(x)].any() and not g[:, idx(x)].any()


def test_dp_ssrr_matches_reference_cell_by_cell(have_ref):
    """The SS/residue DP fill (ExecDPSSwBtck128xKernel<true>) and its walk:
    backtrace codes and matched positions identical to the reference's own
    code, on random SS strings and residue codes that include non-standard
    letters, lower case and gaps (which score 0)."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    rng = np.random.default_rng(21)
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWYXBZUaq-", np.uint8)
    ssl = np.frombuffer(b"HE ", np.uint8)
    for k in range(12):
        Lq, Lr = (int(v) for v in rng.integers(1, 260, 2))
        qss, rss = bytes(rng.choice(ssl, Lq)), bytes(rng.choice(ssl, Lr))
        qr, rr = bytes(rng.choice(letters, Lq)), bytes(rng.choice(letters, Lr))
        bo, po = oracle.dp_ssrr(qss, qr, rss, rr)
        br, pr = oracle.ref_dp_ssrr(qss, qr, rss, rr)
        assert np.array_equal(bo, br), (k, Lq, Lr)
        assert np.array_equal(po, pr), (k, Lq, Lr)


def test_dp_ssrr_identical_sequences_align_diagonally():
    """Equal SS and residue strings: every residue pairs with itself (each
    diagonal step scores >= 1 + 0.2 x a positive self-score, any gap costs 1)."""
    rng = np.random.default_rng(4)
    s = bytes(rng.choice(np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", np.uint8), 90))
    ss = bytes(rng.choice(np.frombuffer(b"HE ", np.uint8), 90))
    _, pos = oracle.dp_ssrr(ss, s, ss, s)
    assert np.array_equal(pos[:, 0], np.arange(89, -1, -1))
    assert np.array_equal(pos[:, 1], pos[:, 0])


@pytest.mark.parametrize("cfg,scale,qsel", [(0, 0.25, None), (1, 0.004, [0, 9])],
                         ids=["c0_1x64", "c1_2x200"])
def test_stage_search_by_ss_matches_reference(have_ref, cfg, scale, qsel):
    """--add-search-by-ss: the oracle's whole stage equals the reference's CPU
    path (with MpStageSSRR between stages 1 and 2) bit for bit."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    qc, rc, _ = synth.config(cfg, scale=scale)
    if qsel:
        qc = qc.subset(qsel)
    opts = default_options(addsearchbyss=1)
    rc = oracle.kept_first(qc, rc, opts)
    o = oracle.process_block(qc, rc, opts)
    r = oracle.ref_process_block(qc, rc, opts, nthreads=1)
    assert_kept_pairs_equal(o, r)


# ---- the whole stage ------------------------------------------------------------------

def test_self_alignment_scores_one():
    """A chain against a rigidly moved copy of itself: TM-score 1, RMSD ~0,
    every residue aligned with itself."""
    rng = np.random.default_rng(8)
    from gtalign_alpha_amd.chunk import Chunk
    c = synth.chain(31, 120)
    R, t = random_rotation(rng), rng.normal(scale=20.0, size=3)
    res = synth.residues(rng, 120)
    q = Chunk.from_structures([c], [res])
    r = Chunk.from_structures([c @ R.T + t], [res])
    o = oracle.process_block(q, r)
    a = o.alndata[0, 0]
    assert abs(a[9] - 1.0) < 1e-4 and abs(a[10] - 1.0) < 1e-4
    # the reference's FP32 RMSD (Kabsch E0 identity) cancels ~1e3 A^2 variances:
    # an exact copy reads a few hundredths of an angstrom, not 0
    assert a[8] < 0.1
    assert a[4] == 120 and a[6] == 120 and a[7] == 0
    tf = o.tfms[0, 0]
    assert np.allclose(tfm_apply(tf, c), c @ R.T + t, atol=2e-2)


def assert_kept_pairs_equal(o, r):
    """Same pairs kept by the stage-1 filter; for those, records, transforms
    and alignment rows bit for bit; the dropped pairs carry no alignment."""
    assert np.array_equal(o.passed, r.passed)
    assert np.array_equal(o.alns, r.alns)
    assert np.array_equal(o.alndata[r.passed], r.alndata[r.passed])
    assert np.array_equal(o.tfms[r.passed], r.tfms[r.passed])
    assert (o.alndata[~r.passed][..., 8] >= 9999).all()


@pytest.mark.parametrize("cfg,scale,qsel", [(0, 0.25, None), (1, 0.004, [0, 9])],
                         ids=["c0_1x64", "c1_2x200"])
def test_stage_matches_reference_cpu_path(have_ref, cfg, scale, qsel):
    """The oracle reproduces the reference's own CPU path (1 thread, with its
    packing of the references kept after stage 1) bit for bit, on references
    ordered so that packing leaves the kept ones in place."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    qc, rc, _ = synth.config(cfg, scale=scale)
    if qsel:
        qc = qc.subset(qsel)
    rc = oracle.kept_first(qc, rc)
    o = oracle.process_block(qc, rc)
    r = oracle.ref_process_block(qc, rc, nthreads=1)
    assert_kept_pairs_equal(o, r)
    assert (r.passed & (o.alndata[..., 8] < 9999)).sum() >= 10


OPTION_PATHS = [dict(refinement=0), dict(refinement=1), dict(refinement=3), dict(refinement=4), dict(depth=0),
                dict(depth=1), dict(depth=3), dict(symmetric=1), dict(referenced=1), dict(nodeletions=1),
                dict(convergence=5), dict(prescore=0.0, scorethld=0.0)]
OPTION_IDS = ["log", "coarse", "full", "fulla", "deep", "high", "shallow", "symmetric", "referenced",
              "nodeletions", "convergence5", "nothresholds"]


@pytest.mark.parametrize("kw", OPTION_PATHS, ids=OPTION_IDS)
def test_option_paths_match_reference_cpu_path(have_ref, kw):
    """Every option path the stage takes (--refinement LOG/COARSE/FULL, --depth
    DEEP/HIGH/SHALLOW, --symmetric, --referenced, --no-deletions,
    --convergence, no score thresholds; CLOptions.cpp:47-48,68-70): the
    oracle equals the reference's own CPU path bit for bit.  The reference
    reads its options once per process, so each option set runs in a child
    process of its own (oracle/ref_cli.py)."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    from oracle import ref_cli
    qc, rc, _ = synth.config(1, scale=0.0015)
    qc = qc.subset([0, 9])            # queries of 574 and 171 residues: both sides longer
    opts = default_options(**kw)
    rc = oracle.kept_first(qc, rc, opts)
    o = oracle.process_block(qc, rc, opts)
    d = ref_cli.run(qc, rc, 1, kw)
    r = oracle.BlockResult(qc.nstrs, rc.nstrs, d["alndata"], d["tfms"], d["alns"], passed=d["passed"])
    aligned = r.passed & (o.alndata[..., 8] < 9999)
    if kw.get("refinement", 2) >= 3:
        # the reference's production full search also runs slots at negative
        # alignment positions, on memory past the pair's alignment (older
        # alignments, the next reference's region): those pairs are left out
        # of the bit-level comparison; at least half of the aligned pairs are
        # compared
        neg = oracle.last_prod_negative(qc.nstrs, rc.nstrs)
        assert (aligned & ~neg).sum() >= 0.5 * aligned.sum()
        assert np.array_equal(o.passed, r.passed)
        assert np.array_equal(o.alns, r.alns)      # rows come before the production search
        m = r.passed & ~neg
        assert np.array_equal(o.alndata[m], r.alndata[m])
        assert np.array_equal(o.tfms[m], r.tfms[m])
        return
    assert_kept_pairs_equal(o, r)
    assert (r.passed & (o.alndata[..., 8] < 9999)).sum() >= 5


def test_reference_filter_blocks_match_reference(have_ref):
    """The filter after stage 1 in blocks of 2 queries (the reference's
    scheduler): a reference stays for both queries of a block when either
    keeps it, and the pairs of dropped references report nothing.  On an
    interleaved chunk (packing moves kept references to new indices):
    - kept sets identical to the reference's;
    - with the reference driver also moving each kept pair's per-slot
      transforms and records when it packs (a diagnostic switch), every kept
      pair's records, transforms and alignment rows are bit-identical;
    - without it, the reference reads per-slot buffers that packing leaves
      behind (another pair's data: MpReform.cpp:57-124 moves only the
      per-pair records, and the per-slot rows change stride with the
      reference count), so some kept pairs differ, mostly in the transform of
      pairs left without an alignment; of the kept pairs with an alignment at
      least 90% are still bit-identical."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    qc, rc, _ = synth.config(1, scale=0.004)
    qc = qc.subset([4, 5, 6, 7])
    o = oracle.process_block(qc, rc, default_options(qgroup=2))
    prev = oracle.ref_set_move_slots(True)
    try:
        moved = oracle.ref_process_block(qc, rc, nthreads=1)
    finally:
        oracle.ref_set_move_slots(prev)
    assert not moved.passed.all() and moved.passed.any()
    assert_kept_pairs_equal(o, moved)
    r = oracle.ref_process_block(qc, rc, nthreads=1)
    assert np.array_equal(o.passed, r.passed)
    same = (o.alndata == r.alndata).all(-1) & (o.tfms == r.tfms).all(-1)
    al = r.passed & (r.alndata[..., 8] < 9999)
    assert al.sum() >= 10 and (same & al).sum() >= 0.9 * al.sum()
    # blocks matter: a pair flagged on its own can continue with its block
    own = oracle.process_block(qc, rc, default_options(qgroup=1))
    assert (o.passed & ~own.passed).any()


def test_options_change_outcome():
    qc, rc, _ = synth.config(0, scale=0.1)
    a = oracle.process_block(qc, rc, default_options())
    b = oracle.process_block(qc, rc, default_options(scorethld=0.0, prescore=0.0))
    # without the score thresholds every pair gets an alignment
    assert (b.alndata[..., 8] < 9999).sum() == qc.nstrs * rc.nstrs
    assert (a.alndata[..., 8] < 9999).sum() < qc.nstrs * rc.nstrs


def test_short_chains_match_reference(have_ref):
    """Chains of 3-8 residues (below the fragment-length floors) and an exact
    copy: the oracle equals the reference's CPU path."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    from gtalign_alpha_amd.chunk import Chunk
    rng = np.random.default_rng(3)
    qs = [synth.chain(1, 3), synth.chain(2, 5), synth.chain(3, 30)]
    rs = [synth.chain(4, 3), synth.chain(5, 4), synth.chain(6, 8), synth.chain(7, 25), qs[2] + 1.0]
    q = Chunk.from_structures(qs, [synth.residues(rng, len(s)) for s in qs])
    r = Chunk.from_structures(rs, [synth.residues(rng, len(s)) for s in rs])
    o = oracle.process_block(q, r, default_options(qgroup=2))
    prev = oracle.ref_set_move_slots(True)
    try:
        ref = oracle.ref_process_block(q, r)
    finally:
        oracle.ref_set_move_slots(prev)
    assert_kept_pairs_equal(o, ref)
    # the 3-residue pair aligns completely; the exact copy scores 1
    assert o.alndata[0, 0, 4] == 30 and abs(o.alndata[0, 0, 9] - 1.0) < 1e-6


# ---- secondary structure (MpSecStr.h:50-56) ------------------------------------------

def na_test_chunk():
    """Nucleic-acid C3' traces (two stem-loops, a compact random trace, a
    201-residue construct with an exact tie) and one protein chain."""
    from gtalign_alpha_amd.chunk import Chunk
    rng = np.random.default_rng(11)
    h1, b1 = synth.na_hairpin(rng, 12, 5)
    h2, b2 = synth.na_hairpin(rng, 30, 7, noise=0.6)
    cmp_ = synth.na_compact(41, 150)
    # tie: residue 0 (G) has two complementary partners at exactly 13.75 A,
    # residues 5 and 130 (C); the lower index, 5, is recorded
    tie = np.stack([100.0 * np.arange(201), np.zeros(201), np.zeros(201)], 1).astype(np.float32)
    tie[0] = (0.0, 0.0, 0.0)
    tie[5] = (13.75, 0.0, 0.0)
    tie[130] = (0.0, 13.75, 0.0)
    tb = ["A"] * 201
    tb[0], tb[5], tb[130] = "G", "C", "C"
    prot = synth.chain(5, 90)
    structs = [h1, h2, cmp_, tie, prot]
    res = [b1, b2, synth.na_bases(rng, 150), "".join(tb), synth.residues(rng, 90)]
    c = Chunk.from_structures(structs, res)
    return synth.mark_nucleic(c, [c.order[i] != 4 for i in range(c.nstrs)])


def na_pairs_brute_force(c, i):
    """Base-pairing letters of nucleic-acid structure i by brute force over
    all position pairs (numpy FP32): partner = the complementary residue more
    than 2 positions away whose C3' distance deviates least from 13.75 A
    (at most 1.25 A), lowest index on ties; mutual partners y < x give '('
    at y and ')' at x."""
    X = c.coords(i).astype(np.float32)
    r = np.frombuffer(c.residues(i).encode(), np.uint8)
    L = len(r)
    d = X[:, None, :] - X[None, :, :]
    dist = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    A, C_, G, T, U = (ord(ch) for ch in "ACGTU")
    a, b = r[:, None], r[None, :]
    comp = (((a == T) | (a == U)) & (b == A)) | ((a == A) & ((b == T) | (b == U))) | \
           ((a == G) & ((b == C_) | (b == U))) | (((a == C_) | (a == U)) & (b == G))
    idx = np.arange(L)
    ok = comp & (np.abs(idx[:, None] - idx[None, :]) > 2)
    dev = np.abs(dist - np.float32(13.75))
    dev = np.where(ok & (dev <= 1.25), dev, np.float32(np.inf))
    part = np.where(np.isfinite(dev.min(1)), dev.argmin(1), -1)
    out = np.full(L, ord("."), np.uint8)
    for y in range(L):
        p = part[y]
        if p > y and part[p] == y:
            out[y], out[p] = ord("("), ord(")")
    return out


def test_na_secstr_is_the_least_deviation_pairing():
    """The oracle's base-pairing SS (the stage's definition) equals the brute
    force above; protein chains keep the C-alpha rule."""
    c = na_test_chunk()
    o = oracle.secstr(c)
    k = {int(c.order[i]): i for i in range(c.nstrs)}
    for i in range(c.nstrs):
        d, l = int(c.dst[i]), int(c.len[i])
        if c.order[i] == 4:
            prot = o[d:d + l].tobytes()
            assert set(prot) <= set(b" het") and b"." not in prot
            continue
        assert np.array_equal(o[d:d + l], na_pairs_brute_force(c, i)), c.order[i]
    tie = o[c.dst[k[3]]:c.dst[k[3]] + 201].tobytes()
    assert tie[0:1] == b"(" and tie[5:6] == b")" and tie[130:131] == b"."
    stem = o[c.dst[k[1]]:c.dst[k[1]] + c.len[k[1]]].tobytes()
    assert stem.count(b"(") >= 25 and stem.count(b"(") == stem.count(b")")


def test_na_secstr_cpu_twin_matches_reference(have_ref):
    """The reference's CPU twin as compiled records for each residue the
    index left by its vectorised minimum (see oracle.set_na_ss_mode); with
    that one step switched, the oracle's SS equals the reference's MpSecStr
    letter for letter and partner for partner, which pins the shared parts
    (pairing rule, distances, deviation bounds, mutual check)."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    c = na_test_chunk()
    prev = oracle.set_na_ss_mode(1)
    try:
        o = oracle.secstr(c)
    finally:
        oracle.set_na_ss_mode(prev)
    r, rpos = oracle.ref_secstr(c, with_pairs=True)
    assert np.array_equal(o, r), np.nonzero(o != r)[0][:10]
    # the compiled minimum loses pairs the least-deviation rule finds (69 vs 48 here)
    assert (oracle.secstr(c) == ord("(")).sum() > (r == ord("(")).sum()


def test_na_stage_matches_reference(have_ref):
    """The whole stage on nucleic-acid chains (SS-weighted stage-2 DP over
    base-pairing SS, nucleic-acid d0 of the final stage): bit for bit the
    reference's CPU path, with the oracle's SS step in the CPU twin's mode."""
    if not have_ref:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    q, r = na_stage_chunks()
    opts = default_options(qgroup=2)
    prev = oracle.set_na_ss_mode(1)
    try:
        o = oracle.process_block(q, r, opts)
    finally:
        oracle.set_na_ss_mode(prev)
    prev = oracle.ref_set_move_slots(True)
    try:
        ref = oracle.ref_process_block(q, r, opts)
    finally:
        oracle.ref_set_move_slots(prev)
    assert_kept_pairs_equal(o, ref)
    # the copy aligns; the protein reference never does
    k = {int(r.order[i]): i for i in range(r.nstrs)}
    qk = {int(q.order[i]): i for i in range(q.nstrs)}
    assert o.alndata[qk[0], k[0], 9] > 0.5
    assert (o.alndata[:, k[3], 8] >= 9999).all()


def na_stage_chunks():
    from gtalign_alpha_amd.chunk import Chunk
    rng = np.random.default_rng(12)
    h, b = synth.na_hairpin(rng, 25, 6)
    qs = [h, synth.na_compact(51, 80)]
    qb = [b, synth.na_bases(rng, 80)]
    R, t = random_rotation(rng), rng.normal(scale=20.0, size=3)
    rs = [(h @ R.T + t + rng.normal(scale=0.5, size=h.shape)).astype(np.float32), synth.na_compact(52, 70),
          synth.na_hairpin(rng, 18, 4)[0], synth.chain(53, 60)]
    rb = [b, synth.na_bases(rng, 70), synth.na_bases(rng, 40), synth.residues(rng, 60)]
    q = Chunk.from_structures(qs, qb)
    r = Chunk.from_structures(rs, rb)
    q = synth.mark_nucleic(q, [True] * q.nstrs)
    r = synth.mark_nucleic(r, [r.order[i] != 3 for i in range(r.nstrs)])
    return q, r
