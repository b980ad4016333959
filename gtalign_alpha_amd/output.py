"""Result writers: the stage's three result sections (alignment records,
transforms, alignment rows) to the reference's plain-text and JSON hit
records.

Follows TdFinalizer::CompressResultsPlain / CompressResultsJSON
(src/libgenp/goutp/TdFinalizerPlain.cpp:31-437, TdFinalizerJSON.cpp:31-387)
and the description formatters of src/libutil/format.h:35-151, under the
reference's default output options (CLOptions.cpp:42-49): --2tm-score 0,
--sort 0 (the greater of the two TM-scores), --nhits 2000, --nalns 2000,
--wrap 80.  Per query, a reference is reported when the greater of its two
TM-scores is at least the score threshold (-s) and its alignment is
non-empty; reports are ordered by descending sort score, stably in
reference order (the reference sorts its record indices by score).

Every field's text is what the reference's printf formats give: the plain
annotation line (serial number in 6 columns, the description in 40 columns
or its last 37 characters after "...", TM-scores %6.4f, RMSD %5.2f, aligned
pairs and boundaries), the alignment record (serial number line, '>' and the
description, lengths and type, scores, identities/matched/gaps with integer
percentages, the five alignment rows wrapped at 80 columns with running
residue numbers, the rotation and translation) and the JSON summary entries
and hit records with their field names and precisions.
"""
import numpy as np

from .capi import (ALN_ALNLENGTH, ALN_BEGCOORDS, ALN_D0Q, ALN_D0R, ALN_ENDCOORDS, ALN_IDNTS, ALN_NGAPS,
                   ALN_PSTVS, ALN_RMSD, ALN_SCOREQ, ALN_SCORER, ALN_2SCOREQ, ALN_2SCORER)

ANNOTATION_DESCLEN = 40           # mylimits.h:51
ANNOTATION_DESCLEN_3L = 37        # mylimits.h:53
DEFAULT_DESCRIPTION_LENGTH = 128  # mylimits.h:62
MAX_DESCRIPTION_LENGTH = 4096     # mylimits.h:65
OUTPUT_INDENT = 13                # mylimits.h:45
NL = "\n"

# sort keys (CLOptions.h TOSortBy, default 0)
SORT_TM_GREATER, SORT_TM_REFERENCE, SORT_TM_QUERY, SORT_TM_HARMONIC, SORT_RMSD = range(5)


def _coords(word):
    """Packed 1-based (query, reference) coordinates of a begin/end field
    (GetCoordY, GetCoordX: low and high 16 bits)."""
    w = int(np.float32(word).view(np.uint32))
    return w & 0xffff, w >> 16


def molecule_type_str(ins_ch_ord0):
    """GetMoleculeTypeStr of a structure's first type word (positive: NA)."""
    return "NA" if ins_ch_ord0 > 0 else "Protein"


def sort_score(a, sortby=SORT_TM_GREATER):
    tq, tr, rmsd = float(a[ALN_SCOREQ]), float(a[ALN_SCORER]), float(a[ALN_RMSD])
    tq32, tr32 = np.float32(tq), np.float32(tr)
    grt = max(tq32, tr32)
    if sortby == SORT_TM_REFERENCE:
        return tr32
    if sortby == SORT_TM_QUERY:
        return tq32
    if sortby == SORT_TM_HARMONIC:
        return np.float32(2.0) * tq32 * tr32 / (tq32 + tr32) if grt > 0 else np.float32(0.0)
    if sortby == SORT_RMSD:
        return np.float32(-rmsd)
    return grt


def format_description(desc, maxoutlen=DEFAULT_DESCRIPTION_LENGTH, indent=OUTPUT_INDENT,
                       width=MAX_DESCRIPTION_LENGTH, addsep=">"):
    """FormatDescription (format.h:35-96): leading blanks dropped, runs of
    blanks/tabs become one blank (or a line break with indentation where the
    next word would pass `width`), at most maxoutlen characters, the last
    three replaced by '...' when cut."""
    out = []
    outpos = linepos = 0
    d = desc.lstrip(" \t\n\r")
    if addsep:
        out.append(addsep)
        outpos += 1
        linepos += 1
    i = 0
    while i < len(d) and outpos < maxoutlen:
        if d[i] in " \t":
            while i < len(d) and d[i] in " \t":
                i += 1
            if width < maxoutlen:
                loc = linepos + 1
                j = i
                while j < len(d) and d[j] not in " \t" and loc < width:
                    j += 1
                    loc += 1
                if width <= loc:
                    out.append(NL + " " * indent)
                    outpos += len(NL) + indent
                    linepos = indent
                    continue
            out.append(" ")
            outpos += 1
            linepos += 1
            continue
        out.append(d[i])
        i += 1
        outpos += 1
        linepos += 1
    s = "".join(out)
    if maxoutlen < outpos and 3 < maxoutlen:
        s = s[:-3] + "..."
    return s


def format_description_json(desc, maxoutlen):
    """FormatDescriptionJSON (format.h:107-151): leading blanks dropped, '"',
    '\\' and '/' escaped, the last maxoutlen characters kept, the first three
    (and escape characters after them) replaced by '.' when cut."""
    d = desc.lstrip(" \t\n\r")
    if not d or maxoutlen <= 0:
        return ""
    esc = "".join("\\" + ch if ch in '"\\/' else ch for ch in d)
    if len(esc) <= maxoutlen:
        return esc
    s = list(esc[len(esc) - maxoutlen:])
    if 3 < maxoutlen:
        s[0] = s[1] = s[2] = "."
        i = 3
        while i < maxoutlen and s[i] in '"\\/':
            s[i] = "."
            i += 1
    return "".join(s)


class Hit:
    """One reported (query, reference) pair of a result."""

    def __init__(self, res, q, r, qi, ri, rdesc):
        self.a = res.alndata[qi, ri]
        self.tfm = res.tfms[qi, ri]
        self.qlen = int(q.len[qi])
        self.rlen = int(r.len[ri])
        self.rdesc = rdesc
        self.rtype = molecule_type_str(int(r.ins_ch_ord[r.dst[ri]]))
        self.alnlen = int(self.a[ALN_ALNLENGTH])
        alns = res.alns if isinstance(res.alns, np.ndarray) else np.frombuffer(res.alns, np.uint8)
        nr = r.nstrs
        off = (qi * r.nposs + nr * (int(q.dst[qi]) + qi)) * 5
        rowlen = r.nposs + nr * (self.qlen + 1)
        beg = off + int(r.dst[ri]) + ri * (self.qlen + 1)
        self.rows = [bytes(alns[beg + rowlen * f: beg + rowlen * f + self.alnlen]).decode("latin-1")
                     for f in range(5)]


def hits(res, q, r, qi, rdescs, scorethld=0.5, sortby=SORT_TM_GREATER):
    """Reported hits of query qi in report order (CompressResultsPlain's
    selection and TdAlnWriter's sort by descending score, stable)."""
    out = []
    for ri in range(r.nstrs):
        a = res.alndata[qi, ri]
        grt = max(np.float32(a[ALN_SCOREQ]), np.float32(a[ALN_SCORER]))
        if grt < np.float32(scorethld) or int(a[ALN_ALNLENGTH]) < 1:
            continue
        out.append((sort_score(a, sortby), ri))
    order = sorted(range(len(out)), key=lambda k: -out[k][0])
    return [Hit(res, q, r, qi, out[k][1], rdescs[out[k][1]]) for k in order]


def annotation_plain(h, serial):
    """MakeAnnotationPlain (TdFinalizerPlain.cpp:179-221) with the serial
    number written into its 6 reserved columns."""
    a = h.a
    qb, tb = _coords(a[ALN_BEGCOORDS])
    qe, te = _coords(a[ALN_ENDCOORDS])
    d = h.rdesc
    s = "%6d" % serial
    if len(d) <= ANNOTATION_DESCLEN:
        s += " %-40s" % d
    else:
        s += " ..." + "%-37s" % d[len(d) - ANNOTATION_DESCLEN_3L:]
    s += " %6.4f %6.4f %5.2f" % (a[ALN_SCORER], a[ALN_SCOREQ], a[ALN_RMSD])
    s += " %5d %5d-%-5d %5d-%-5d %5d%s" % (h.alnlen - int(a[ALN_NGAPS]), qb, qe, tb, te, h.rlen, NL)
    return s


def alignment_plain(h, serial, wrap=80, referenced=False, nodeletions=False, second_tm=False):
    """The alignment record of CompressResultsPlain (serial line, description,
    FormatScoresPlain, FormatAlignmentPlain, FormatFooterPlain;
    TdFinalizerPlain.cpp:134-156, 228-437)."""
    a, n = h.a, h.alnlen
    out = [("%d." % serial).ljust(13) + NL, format_description(h.rdesc) + NL]
    out.append("  Length: Refn. = %d, Query = %d   Type:%s%s%s" % (h.rlen, h.qlen, h.rtype, NL, NL))
    out.append(" TM-score (Refn./Query) = %.5f / %.5f, d0 (Refn./Query) = %.2f / %.2f,  RMSD = %.2f A%s" %
               (a[ALN_SCORER], a[ALN_SCOREQ], a[ALN_D0R], a[ALN_D0Q], a[ALN_RMSD], NL))
    if second_tm:
        out.append(" 2TM-score (Refn./Query) = %.5f / %.5f%s" % (a[ALN_2SCORER], a[ALN_2SCOREQ], NL))
    idts, psts, gaps = int(a[ALN_IDNTS]), int(a[ALN_PSTVS]), int(a[ALN_NGAPS])
    line = ""
    if idts > 0:
        line += " Identities = %d/%d (%d%%)" % (idts, n, idts * 100 // n)
    if psts > 0:
        line += ("," if idts else "") + " Matched = %d/%d (%d%%)" % (psts, n, psts * 100 // n)
    if gaps > 0:
        line += ("," if idts or psts else "") + " Gaps = %d/%d (%d%%)" % (gaps, n, gaps * 100 // n)
    out.append(line + NL + NL)
    qb, tb = _coords(a[ALN_BEGCOORDS])
    _, te = _coords(a[ALN_ENDCOORDS])
    qrow, mid, trow, qss, tss = h.rows
    for f in range(0, n, wrap):
        nb = min(wrap, n - f)
        out.append("%-13s" % "struct" + qss[f:f + nb] + NL)
        seg = qrow[f:f + nb]
        g = seg.count("-")
        qb += nb - g
        out.append("Query: %5u " % (qb - (nb - g)) + seg + " %-5d%s" % (qb if nb <= g else qb - 1, NL))
        out.append(" " * 13 + mid[f:f + nb] + NL)
        seg = trow[f:f + nb]
        g = seg.count("-")
        head = "Refn.: %5s " % "..." if (nodeletions and f) else "Refn.: %5u " % tb
        tb += nb - g
        if nodeletions:
            tail = " %-5d%s" % (te, NL) if n <= f + wrap else " %-5s%s" % ("...", NL)
        else:
            tail = " %-5d%s" % (tb if nb <= g else tb - 1, NL)
        out.append(head + seg + tail)
        out.append("%-13s" % "struct" + tss[f:f + nb] + NL)
        out.append(NL)
    t = h.tfm
    out.append(" Rotation [3,3] and translation [3,1] for %s:%s" % ("Reference" if referenced else "Query", NL))
    for i in range(3):
        out.append("  %10.6f %10.6f %10.6f    %14.6f%s" % (t[3 * i], t[3 * i + 1], t[3 * i + 2], t[9 + i], NL))
    out.append(NL)
    out.append(NL + NL)
    return "".join(out)


def summary_json(h):
    """MakeAnnotationJSON (TdFinalizerJSON.cpp:169-217)."""
    a = h.a
    qb, tb = _coords(a[ALN_BEGCOORDS])
    qe, te = _coords(a[ALN_ENDCOORDS])
    return ('    {"summary_entry": {' + NL +
            '      "description": "' + format_description_json(h.rdesc, ANNOTATION_DESCLEN) + '",' + NL +
            '      "tmscore_refn": %.4f,%s' % (a[ALN_SCORER], NL) +
            '      "tmscore_query": %.4f,%s' % (a[ALN_SCOREQ], NL) +
            '      "rmsd": %.2f,%s' % (a[ALN_RMSD], NL) +
            '      "n_aligned": %d,%s' % (h.alnlen - int(a[ALN_NGAPS]), NL) +
            '      "query_from": %d,%s' % (qb, NL) +
            '      "query_to": %d,%s' % (qe, NL) +
            '      "refn_from": %d,%s' % (tb, NL) +
            '      "refn_to": %d,%s' % (te, NL) +
            '      "refn_length": %d%s' % (h.rlen, NL) +
            '    }}')


def hit_record_json(h, referenced=False, second_tm=False):
    """The JSON hit record of CompressResultsJSON (TdFinalizerJSON.cpp:131-160,
    224-387)."""
    a, n = h.a, h.alnlen
    s = ('    {"hit_record": {' + NL + '      "reference_description": "' +
         format_description_json(h.rdesc, DEFAULT_DESCRIPTION_LENGTH) + '",' + NL)
    s += ('      "query_length": %d,%s      "reference_length": %d,%s      "alignment_type": "%s",%s'
          '      "alignment": {%s' % (h.qlen, NL, h.rlen, NL, h.rtype, NL, NL))
    tm2q = tm2r = "NA"
    if second_tm:
        tm2q, tm2r = "%.5g" % a[ALN_2SCOREQ], "%.5g" % a[ALN_2SCORER]
    s += ('        "tmscore_refn": %.5f,%s        "tmscore_query": %.5f,%s        "d0_refn": %.2f,%s'
          '        "d0_query": %.2f,%s        "rmsd": %.2f,%s        "2tmscore_refn": "%s",%s'
          '        "2tmscore_query": "%s",%s        "n_identities": %d,%s        "n_matched": %d,%s'
          '        "n_gaps": %d,%s        "n_aligned": %d,%s' %
          (a[ALN_SCORER], NL, a[ALN_SCOREQ], NL, a[ALN_D0R], NL, a[ALN_D0Q], NL, a[ALN_RMSD], NL, tm2r, NL,
           tm2q, NL, int(a[ALN_IDNTS]), NL, int(a[ALN_PSTVS]), NL, int(a[ALN_NGAPS]), NL, n, NL))
    qb, tb = _coords(a[ALN_BEGCOORDS])
    qrow, mid, trow, qss, tss = h.rows
    gq, gr = qrow.count("-"), trow.count("-")
    qe, re_ = qb + n - gq, tb + n - gr
    s += ('        "query_from": %u,%s        "query_to": %u,%s        "refn_from": %u,%s'
          '        "refn_to": %u,%s' % (qb, NL, qe if n <= gq else qe - 1, NL, tb, NL, re_ if n <= gr else re_ - 1, NL))
    s += '        "query_secstr": "' + qss + '",' + NL + '        "refrn_secstr": "' + tss + '",' + NL
    s += '        "query_aln": "' + qrow + '",' + NL + '        "refrn_aln": "' + trow + '",' + NL
    s += '        "middle": "' + mid + '"' + NL
    s += "      }," + NL
    t = h.tfm
    s += '      "tfm_referenced": %d,%s' % (1 if referenced else 0, NL)
    s += ('      "rotation_matrix_rowmajor": [' + ", ".join("%.6f" % t[k] for k in range(9)) + "]," + NL)
    s += '      "translation_vector": [%.6f, %.6f, %.6f]%s' % (t[9], t[10], t[11], NL)
    s += "    }}"
    return s


def query_section_plain(res, q, r, qi, rdescs, scorethld=0.5, nhits=2000, nalns=2000, **kw):
    """The per-query body of a plain output file: the numbered annotation
    lines of the top `nhits` hits, then the alignment records of the top
    `nalns` hits."""
    hs = hits(res, q, r, qi, rdescs, scorethld)
    ann = "".join(annotation_plain(h, k + 1) for k, h in enumerate(hs[:nhits]))
    alns = "".join(alignment_plain(h, k + 1, **kw) for k, h in enumerate(hs[:nalns]))
    return ann, alns


def query_section_json(res, q, r, qi, rdescs, scorethld=0.5, nhits=2000, nalns=2000, **kw):
    """The per-query JSON lists: summary entries and hit records."""
    hs = hits(res, q, r, qi, rdescs, scorethld)
    return ([summary_json(h) for h in hs[:nhits]], [hit_record_json(h, **kw) for h in hs[:nalns]])
