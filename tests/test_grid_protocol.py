"""CPU model of the device-wide DP fill's strip protocol (k_dp_grid).

The fill hands (pair, strip) items to one-wave workgroups from a counter in
increasing order; strip s of a pair reads strip s-1's last row after waiting
for the number of columns strip s-1 has published.  This model replays every
wait, row read, row write
and publish that k_dp_grid and dp_strip_impl perform, in their order, with
generators standing in for the workgroups:

- item loop and pair lookup: k_dp_grid (gta_kernels.hip), binary search over
  the first strip item of each pair;
- strip s > 0 first waits for min(128, Lr) columns and reads columns
  0..min(128, Lr)-1 (the ring's first block and the block waiting in
  registers);
- at each step t with t mod 64 == 63 it waits, when t + 65 < Lr, for
  min(t + 129, Lr) columns and reads columns t+65 .. t+128 below Lr;
- a strip that stores its row (every strip but the pair's last) writes
  columns jo-(jo mod 64) .. jo at step t = jo + 63 whenever jo mod 64 == 63
  or jo == Lr - 1, then publishes jo + 1;
- after the strip k_dp_grid publishes Lr.

The checks: every read column was written and published before the read,
no schedule of the workers deadlocks, and every item is taken.  The shapes
are those of the GPU tests that use the device-wide form: every pair with
Lq*Lr >= 0 cells, >= 2 strips and Lr >= 256 (the thresholds those tests set,
default_dp_forms in gta_capi.hip), over the SS/residue fill's workloads.
"""
import random
import numpy as np
import pytest

from gtalign_alpha_amd import synth


class StripProtocol:
    """Workers are generators that yield before every wait and after every
    publish; a seeded scheduler picks which runnable worker moves next
    (policy "greedy-reader" starts every worker on an item, then always moves
    the worker on the highest item, so readers run as far ahead of their
    writers as the waits allow).  A wait
    that is not yet satisfied leaves its worker blocked; no runnable worker
    while items remain unfinished is a deadlock."""

    def __init__(self, pairs, wait_ahead=129, skip=(), skip_items=()):
        # pairs: (Lq, Lr) of the grid bucket in order; skip: pair indices whose
        # dp_setup fails (all their strips are dropped without a publish);
        # skip_items: single strips dropped (a planted mistake)
        self.pairs = pairs
        self.skip_items = set(skip_items)
        self.wait_ahead = wait_ahead
        self.skip = set(skip)
        self.strip0 = [0]
        for Lq, _ in pairs:
            self.strip0.append(self.strip0[-1] + (Lq + 63) // 64)
        self.total = self.strip0[-1]
        self.done = [0] * self.total
        self.written = [None] * self.total
        self.next = 0
        self.taken = 0
        self.errors = []

    def read(self, item, lo, hi):
        # columns lo..hi-1 of the row of item (strip s-1)
        if hi <= lo:
            return
        w = self.written[item]
        if self.done[item] < hi or w is None or not w[lo:hi].all():
            self.errors.append(f"read of columns {lo}..{hi - 1} of item {item} before they were published")

    def strip(self, k, s):
        Lq, Lr = self.pairs[k]
        nstrips = self.strip0[k + 1] - self.strip0[k]
        me = self.strip0[k] + s
        up = me - 1 if s > 0 else None
        putrow = s + 1 < nstrips
        if putrow:
            self.written[me] = np.zeros(Lr, bool)
        if up is not None:
            yield (up, min(128, Lr))
            self.read(up, 0, min(128, Lr))
        for t in range(Lr + 63):
            if (t & 63) == 63 and up is not None:
                if t + 65 < Lr:
                    yield (up, min(t + self.wait_ahead, Lr))
                self.read(up, t + 65, min(t + 129, Lr))
            jo = t - 63
            if putrow and jo >= 0 and ((jo & 63) == 63 or jo == Lr - 1):
                self.written[me][jo - (jo & 63):jo + 1] = True
                self.done[me] = max(self.done[me], jo + 1)
                yield None
        self.done[me] = max(self.done[me], Lr)

    def worker(self):
        while True:
            item = self.next
            self.next += 1
            if item >= self.total:
                return
            k = int(np.searchsorted(self.strip0, item, side="right")) - 1
            self.taken += 1
            if k in self.skip or item in self.skip_items:
                continue
            yield ("item", item)
            yield from self.strip(k, item - self.strip0[k])

    def run(self, nworkers, policy="random", seed=0):
        rng = random.Random(seed)
        ws = []                    # [generator, pending wait, current item]
        for _ in range(min(nworkers, self.total)):
            ws.append([self.worker(), None, -1])
        while ws:
            ready = [w for w in ws if w[1] is None or self.done[w[1][0]] >= w[1][1]]
            if not ready:
                self.errors.append("deadlock: every worker waits: " +
                                   ", ".join(f"item {w[2]} on item {w[1][0]} for {w[1][1]} columns" for w in ws))
                return self
            if policy == "greedy-reader":
                w = max(ready, key=lambda x: (x[2] < 0, x[2]))   # idle workers take items first
            else:
                w = rng.choice(ready)
            try:
                ev = next(w[0])
            except StopIteration:
                ws.remove(w)
                continue
            if isinstance(ev, tuple) and ev[0] == "item":
                w[1], w[2] = None, ev[1]
            else:
                w[1] = ev
        return self


def grid_pairs(qlens, rlens, min_cells=0, min_strips=2):
    return [(int(a), int(b)) for a in qlens for b in rlens
            if a * b >= min_cells and (a + 63) // 64 >= min_strips and b >= 256]


def ssrr_workloads():
    """(Lq, Lr) grid pairs of the SS/residue fill's GPU tests: the DP-kernel
    test's 2 x 6 chains of 300-900 residues (seed 12, lengths drawn as the
    test draws them) and the stage test's 3 queries x 30 references of >= 256
    residues."""
    out = {}
    rng = np.random.default_rng(12)
    qlen = [int(rng.integers(300, 900)) for _ in range(2)]
    # the references' lengths change with the deletions and insertions of
    # the noisy copies, so the model sweeps the whole drawn range instead
    out["dp-kernel"] = grid_pairs(qlen, list(range(256, 900, 37)) + [900, 1023, 1024, 1025])
    qc, rc, _ = synth.config(1, scale=0.002)
    q = qc.subset([0, 7, 15])
    rl = [int(v) for v in rc.len if v >= 256][:30]
    out["stage"] = grid_pairs([int(v) for v in q.len], rl)
    return out


@pytest.mark.parametrize("name", ["dp-kernel", "stage"])
@pytest.mark.parametrize("nworkers,policy", [(1, "random"), (3, "random"), (16, "random"), (512, "random"),
                                             (3, "greedy-reader"), (512, "greedy-reader")])
def test_grid_strip_protocol_drains(name, nworkers, policy):
    pairs = ssrr_workloads()[name]
    assert pairs
    m = StripProtocol(pairs).run(nworkers, policy, seed=nworkers)
    assert not m.errors, m.errors[:3]
    assert m.taken == m.total
    assert all(d >= pairs[k][1] for k in range(len(pairs))
               for d in m.done[m.strip0[k]:m.strip0[k + 1]])


def test_grid_strip_protocol_with_dropped_pairs():
    """Pairs whose setup fails drop all their strips; the others still drain."""
    pairs = ssrr_workloads()["stage"]
    m = StripProtocol(pairs, skip=range(0, len(pairs), 3)).run(16, "greedy-reader")
    assert not m.errors, m.errors[:3]
    assert m.taken == m.total


def test_grid_strip_protocol_catches_a_short_wait():
    """A planted mistake in the model (waiting for 64 columns fewer than the
    strip then reads) is reported as a read before publication."""
    pairs = [(200, 700), (130, 1000)]
    m = StripProtocol(pairs, wait_ahead=65).run(4, "greedy-reader")
    assert any("before they were published" in e for e in m.errors)


def test_grid_strip_protocol_catches_a_dropped_strip():
    """A planted mistake in the model (one strip of a pair dropped while the
    strip below it runs) is reported as a deadlock."""
    m = StripProtocol([(200, 700), (130, 1000)], skip_items=[1]).run(4)
    assert any(e.startswith("deadlock") for e in m.errors)
