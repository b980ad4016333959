"""Result writers (gtalign_alpha_amd/output.py) against the reference's own
plain output: the hit section of out/2021-09-18_00000039_1__0.md (two
references of a 1,042-residue query), kept verbatim as a data fixture
(tests/golden/plain_hits_example.txt).  The fixture's records are parsed back
into alignment records, rows and transforms; formatting them must give the
fixture's text byte for byte.  JSON records are checked for structure and
for the values they carry."""
import json
import os
import re
import types

import numpy as np

from gtalign_alpha_amd import output
from gtalign_alpha_amd.capi import (ALN_ALNLENGTH, ALN_BEGCOORDS, ALN_D0Q, ALN_D0R, ALN_ENDCOORDS, ALN_IDNTS,
                                    ALN_NGAPS, ALN_PSTVS, ALN_RMSD, ALN_SCOREQ, ALN_SCORER)

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "plain_hits_example.txt")


def packed(qpos, rpos):
    return np.uint32((rpos << 16) | qpos).view(np.float32)


def parse_record(lines):
    """Hit fields from one plain alignment record of the fixture."""
    desc = lines[1][1:]
    m = re.match(r"  Length: Refn\. = (\d+), Query = (\d+)   Type:(\w+)", lines[2])
    rlen, qlen, rtype = int(m.group(1)), int(m.group(2)), m.group(3)
    m = re.match(r" TM-score \(Refn\./Query\) = ([\d.]+) / ([\d.]+), d0 \(Refn\./Query\) = ([\d.]+) / ([\d.]+),"
                 r"  RMSD = ([\d.]+) A", lines[4])
    tmr, tmq, d0r, d0q, rmsd = (float(g) for g in m.groups())
    stats = dict((k, int(v)) for k, v in re.findall(r"(\w+) = (\d+)/\d+", lines[5]))
    rows = ["", "", "", "", ""]
    k = 7
    qbeg = rbeg = qend = rend = None
    while lines[k].startswith("struct"):
        q, mid, t, qs, ts = lines[k + 1], lines[k + 2], lines[k + 3], lines[k], lines[k + 4]
        num = q.rstrip().split()[-1]
        nb = len(q) - 13 - 1 - max(5, len(num))
        if qbeg is None:
            qbeg, rbeg = int(q[6:12]), int(t[6:12])
        qend, rend = int(num), int(t.rstrip().split()[-1])
        for f, line in ((0, q), (1, mid), (2, t), (3, qs), (4, ts)):
            rows[f] += line[13:13 + nb].ljust(nb)
        k += 6
    tfm = np.zeros(12, np.float32)
    for i in range(3):
        v = [float(x) for x in lines[k + 1 + i].split()]
        tfm[3 * i: 3 * i + 3] = v[:3]
        tfm[9 + i] = v[3]
    a = np.zeros(15, np.float32)
    a[ALN_SCORER], a[ALN_SCOREQ], a[ALN_D0R], a[ALN_D0Q], a[ALN_RMSD] = tmr, tmq, d0r, d0q, rmsd
    a[ALN_IDNTS], a[ALN_PSTVS], a[ALN_NGAPS] = stats.get("Identities", 0), stats.get("Matched", 0), stats.get("Gaps", 0)
    a[ALN_ALNLENGTH] = len(rows[0])
    a[ALN_BEGCOORDS] = packed(qbeg, rbeg)
    a[ALN_ENDCOORDS] = packed(qend, rend)
    return types.SimpleNamespace(a=a, tfm=tfm, qlen=qlen, rlen=rlen, rdesc=desc, rtype=rtype,
                                 alnlen=len(rows[0]), rows=rows)


def fixture():
    text = open(FIXTURE).read()
    lines = text.split("\n")
    ann = [l + "\n" for l in lines[:2]]
    starts = [i for i, l in enumerate(lines) if re.match(r"^\d+\.\s+$", l)]
    recs = []
    for n, s in enumerate(starts):
        e = starts[n + 1] if n + 1 < len(starts) else len(lines) - 1
        recs.append("\n".join(lines[s:e]) + "\n")
    return ann, recs, [parse_record(lines[s:]) for s in starts]


def test_plain_records_reproduce_the_reference_output():
    ann, recs, hs = fixture()
    assert len(hs) == 2 and hs[0].alnlen == 180 and hs[1].alnlen == 453
    for k, h in enumerate(hs):
        assert output.annotation_plain(h, k + 1) == ann[k]
        assert output.alignment_plain(h, k + 1) == recs[k]


def test_description_formatting():
    assert output.format_description("  a  b\tc") == ">a b c"
    long = "x" * 200
    s = output.format_description(long)
    # the copy stops at 128 output characters; '...' replaces the end only
    # when a wrap pushed the count past the limit (format.h:93-95)
    assert s == ">" + "x" * 127
    assert output.format_description_json('a"b/c', 40) == 'a\\"b\\/c'
    assert output.format_description_json("y" * 50, 40) == "..." + "y" * 37


def test_json_records_parse_and_carry_the_values():
    _, _, hs = fixture()
    h = hs[0]
    rec = json.loads(output.hit_record_json(h))["hit_record"]
    assert rec["query_length"] == 1042 and rec["reference_length"] == 81
    al = rec["alignment"]
    assert al["tmscore_refn"] == 0.53386 and al["tmscore_query"] == 0.06122 and al["rmsd"] == 3.52
    assert al["n_identities"] == 5 and al["n_matched"] == 60 and al["n_gaps"] == 110 and al["n_aligned"] == 180
    assert (al["query_from"], al["query_to"], al["refn_from"], al["refn_to"]) == (343, 511, 1, 81)
    assert al["query_aln"] == h.rows[0] and al["middle"] == h.rows[1]
    assert rec["translation_vector"] == [45.347839, -2.546942, 24.362478]
    summ = json.loads(output.summary_json(h))["summary_entry"]
    assert summ["n_aligned"] == 70 and summ["refn_length"] == 81
    assert summ["description"].startswith("...")
