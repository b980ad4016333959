"""PDB and PDBx/mmCIF readers: one structure per file, into the chunk format.

Follows FlexDataRead::ReadDataPDB (src/libgenp/gdats/FlexDataRead.cpp:282-547)
under the reference's default reading options (CLOptions.cpp:51-59): --ter 3
(a TER or END record after the first residue, or a change of chain, ends the
file's one structure), --split 0, --hetatm 0, --mol 0 (molecule type decided
per residue), --aatom " CA ", --natom " C3'".  Per record:

- MODEL records set the model number (non-blank characters of columns 11-14);
- ATOM records (HETATM too with hetatm=True) whose alternate location
  (column 17) is blank or 'A' and whose atom name (columns 13-16) is the
  protein or nucleic-acid atom give one residue each: chain id (column 22),
  insertion code (column 27), residue name (columns 18-20) mapped by
  `residue_code`, residue number (columns 23-26) and x, y, z (columns 31-38,
  39-46, 47-54), parsed like std::stoi / std::stof (leading blanks, the
  longest numeric prefix); a field with no number is an error;
- a residue whose name has a blank first column and a blank or 'D' second
  column counts as nucleic acid, any other as protein; the structure's type
  is the sign of their sum (PM2DVectorFields.h:96-115).

Each residue's insertion/chain/order word is (ins << 16) | (chain << 8) |
order (PM2DVectorFields.h:40-41), negated for a protein residue; the first
residue's word carries the structure's type (PMBatchStrData.h:368-412,
476-477).  A structure shorter than min_length (--dev-min-length, default 20,
CLOptions.cpp:91) or with no selected atom is skipped, as
PMBatchStrData::FinalizeCrntStructure does (PMBatchStrData.h:439-451).
PMBatchStrData::AddOneResidue (PMBatchStrData.h:368-386) abandons a structure
when a residue arrives while more than max_length residues (when max_length
> 0) or more than 65,535 are already stored, so it keeps at most
max_length + 1 (at most 65,536).  The reference reads queries with no limit
(TaskScheduler.cpp:483-488) and references with --dev-max-length (default
4000, CLOptions.cpp:90; TaskScheduler.cpp:490-494): `read_chunk` takes the
limit as an argument, DEV_MAX_LENGTH for a reference chunk and None (no limit)
for a query chunk.  The stage library accepts structures of up to 65,535
residues (the 16-bit packed alignment coordinates of gproc.h:207-227), so a
65,536-residue structure is read but rejected at upload.

Lines shorter than 80 columns are read as if padded with blanks to 80; the
reference reads such lines' missing columns from the bytes that follow.
Files ending in .gz are decompressed.
"""
import gzip
import re
from dataclasses import dataclass

import numpy as np

from .chunk import Chunk

DEFAULT_MIN_LENGTH = 20
DEV_MAX_LENGTH = 4000        # --dev-max-length default, for reference chunks
_BUFFER_LIMIT = 65535        # PMBSmaxonestructurelength (PMBatchStrData.h:34)
PROTEIN_ATOM = " CA "
NUCLEIC_ATOM = " C3'"

# three-letter residue names with a one-letter code (libutil/alpha.cpp:120-143)
_RESIDUE_CODES = {
    "ALA": "A", "DAL": "A", "ASX": "B", "CYS": "C", "DCY": "C", "ASP": "D", "DAS": "D",
    "GLU": "E", "DGL": "E", "PHE": "F", "DPN": "F", "GLY": "G", "HIS": "H", "DHI": "H",
    "ILE": "I", "DIL": "I", "LYS": "K", "DLY": "K", "LEU": "L", "DLE": "L",
    "MET": "M", "MED": "M", "MSE": "M", "ASN": "N", "DSG": "N", "PYL": "O",
    "PRO": "P", "DPR": "P", "GLN": "Q", "DGN": "Q", "ARG": "R", "DAR": "R",
    "SER": "S", "DSN": "S", "THR": "T", "DTH": "T", "SEC": "U", "VAL": "V", "DVA": "V",
    "TRP": "W", "DTR": "W", "TYR": "Y", "DTY": "Y", "GLX": "Z",
}

_INT_PREFIX = re.compile(r"\s*([+-]?\d+)")
_FLOAT_PREFIX = re.compile(r"\s*([+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?))")


def residue_code(name3):
    """One-letter code of a 3-column residue name (libutil/alpha.cpp:65-89): the
    table above, else the name's only non-blank letter when it stands alone in
    column 1, or in column 2 or 3 after blanks, else 'X'."""
    code = _RESIDUE_CODES.get(name3)
    if code is not None:
        return code
    a, b, c = name3[0], name3[1], name3[2]
    if a == " ":
        if b == " ":
            return "X" if c == " " else c
        return b if c == " " else "X"
    return a if b == " " else "X"


def _stoi(s):
    m = _INT_PREFIX.match(s)
    if m is None:
        raise ValueError("no residue number in %r" % s)
    return int(m.group(1))


def _stof(s):
    m = _FLOAT_PREFIX.match(s)
    if m is None:
        raise ValueError("no coordinate in %r" % s)
    return float(np.float32(float(m.group(1))))


@dataclass
class PdbStructure:
    coords: np.ndarray        # (L, 3) float32
    residues: str             # one-letter codes
    resnum: np.ndarray        # (L,) int32 residue numbers as in the file
    ins_ch_ord: np.ndarray    # (L,) int32 insertion/chain/order words, type-marked
    chain: str
    model: str
    nucleic: bool

    def __len__(self):
        return len(self.residues)


def _lines(source):
    if isinstance(source, (bytes, bytearray)):
        return source.decode("latin-1").splitlines()
    if "\n" in source:
        return source.splitlines()
    opener = gzip.open if source.endswith(".gz") else open
    with opener(source, "rb") as f:
        return f.read().decode("latin-1").splitlines()


def read_pdb(source, hetatm=False, min_length=DEFAULT_MIN_LENGTH, max_length=None,
             protein_atom=PROTEIN_ATOM, nucleic_atom=NUCLEIC_ATOM):
    """The file's structure (path, or PDB text containing a newline), or None
    when no atom is selected, the chain is shorter than min_length, or a
    residue arrives while more than max_length (None or <= 0: no limit) or
    more than 65,535 residues are stored."""
    if len(protein_atom) != 4 or len(nucleic_atom) != 4:
        raise ValueError("atom names are 4 columns wide")
    model = ""
    chain = None
    chord = 0
    moltype = 0
    xyz, res, rnum, ico = [], [], [], []
    for line in _lines(source):
        if len(line) < 3:
            continue
        if xyz and (line.startswith("END") or line.startswith("TER")):
            break
        line = line.ljust(80)
        if line.startswith("MODEL "):
            model = "".join(ch for ch in line[10:14] if ch not in " \t")
            continue
        if line[16] not in " A":
            continue
        if not (line.startswith("ATOM  ") or (hetatm and line.startswith("HETATM"))):
            continue
        if line[12:16] != protein_atom and line[12:16] != nucleic_atom:
            continue
        chid, insc = line[21], line[26]
        resd = residue_code(line[17:20])
        resnum = _stoi(line[22:26])
        x, y, z = _stof(line[30:38]), _stof(line[38:46]), _stof(line[46:54])
        if chain != chid:
            if xyz:
                break
            chain = chid
            chord += 1
        restype = 1 if (line[17] == " " and line[18] in " D") else -1
        moltype += restype
        if (max_length is not None and 0 < max_length < len(xyz)) or _BUFFER_LIMIT < len(xyz):
            return None
        word = ((ord(insc) & 0xff) << 16) | ((ord(chid) & 0xff) << 8) | (chord & 0xff)
        xyz.append((x, y, z))
        res.append(resd)
        rnum.append(resnum)
        ico.append(word if restype > 0 else -word)
    if len(xyz) < 1 or len(xyz) < min_length:
        return None
    ico = np.array(ico, dtype=np.int32)
    nucleic = moltype > 0
    ico[0] = abs(int(ico[0])) if nucleic else -abs(int(ico[0]))
    return PdbStructure(np.array(xyz, dtype=np.float32), "".join(res),
                        np.array(rnum, dtype=np.int32), ico, chain, model, nucleic)


def chunk_from_structures(structs):
    """Chunk of PdbStructure records, sorted by descending length (stable), with
    their residue numbers and insertion/chain/order words."""
    c = Chunk.from_structures([s.coords for s in structs], [s.residues for s in structs])
    if structs:
        c.resnum = np.concatenate([structs[i].resnum for i in c.order]).astype(np.int32)
        c.ins_ch_ord = np.concatenate([structs[i].ins_ch_ord for i in c.order]).astype(np.int32)
    return c


def read_chunk(sources, max_length=None, **kw):
    """Chunk of the structures of the given PDB or PDBx/mmCIF files (skipped files are left
    out); returns (chunk, indices of the sources that gave a structure).
    max_length: None for a query chunk (the reference reads queries with no
    limit), DEV_MAX_LENGTH for a reference chunk read as the reference does
    by default."""
    structs, kept = [], []
    for i, src in enumerate(sources):
        s = read_structure(src, max_length=max_length, **kw)
        if s is not None:
            structs.append(s)
            kept.append(i)
    return chunk_from_structures(structs), kept


# ---- PDBx/mmCIF (FlexDataRead::ReadDataCIF, FlexDataRead.cpp:553-908) --------------

_CIF_FIELDS = ("group_PDB", "label_atom_id", "label_alt_id", "label_comp_id", "label_asym_id",
               "label_seq_id", "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z", "auth_seq_id",
               "auth_asym_id", "pdbx_PDB_model_num")


_CIF_LINE_WIDTH = 128


def _cif_field(name):
    """The _atom_site field a header line names (FlexDataRead.cpp:597-618):
    matched by prefix, Cartn_x/y/z only when followed by white space."""
    for f in _CIF_FIELDS:
        if name.startswith(f):
            if f.startswith("Cartn_") and not (len(name) > 7 and name[7] in " \t\r\n"):
                continue
            return f
    return None


def _cif_atom_is(token, atom):
    """lfCheckAtom (FlexDataRead.cpp:627-636): the token, after an optional
    opening double quote, starts with the trimmed atom name followed by a
    double quote or the end of the token."""
    t = token[1:] if token.startswith('"') else token
    if not t.startswith(atom):
        return False
    rest = t[len(atom):]
    return rest == "" or rest[0] == '"'


def read_cif(source, hetatm=False, min_length=DEFAULT_MIN_LENGTH, max_length=None,
             protein_atom=PROTEIN_ATOM, nucleic_atom=NUCLEIC_ATOM):
    """The file's structure from a PDBx/mmCIF file (path, or text containing a
    newline) under the reference's default reading options, as
    FlexDataRead::ReadDataCIF reads it:

    - an _atom_site loop's header lines set the column of each field; a line
      starting with '#' or 'loop_' ends the section; the section is read once
      group_PDB, label_atom_id, label_comp_id, a chain (auth_asym_id or
      label_asym_id), a residue number (auth_seq_id or label_seq_id) and
      Cartn_x/y/z are all known;
    - rows are split at blanks and tabs within their first 128 columns; ATOM rows (HETATM too with
      hetatm=True) whose label_alt_id is '.' or 'A' (when the field exists)
      and whose label_atom_id is the protein or nucleic-acid atom (trimmed,
      optionally double-quoted) give one residue each; rows with fewer tokens
      than a needed column are skipped;
    - chain id: first character of auth_asym_id (label_asym_id without it),
      '.' read as blank; insertion code: first character of
      pdbx_PDB_ins_code, '?' read as blank; residue number: auth_seq_id
      (label_seq_id without it) parsed like std::stoi; coordinates like
      std::stof; residue code: the 3 bytes at the start of label_comp_id
      mapped as for PDB names (so a 1-letter name followed by a blank gives
      its letter, 'DA' gives 'X');
    - a change of pdbx_PDB_model_num or of chain after the first residue
      ends the structure (--ter 3, --split 0);
    - a residue is nucleic acid when label_comp_id has fewer than 2
      characters or starts with 'D', protein otherwise.
    Length limits, type words and the structure's type are as read_pdb's."""
    pa, na = protein_atom.strip(), nucleic_atom.strip()
    cols = {}
    loop = -1
    ready = False
    model_prev = chain_prev = None
    chord = 0
    moltype = 0
    xyz, res, rnum, ico = [], [], [], []
    chain0 = model0 = ""
    for line in _lines(source):
        if len(line) < 5 or line.startswith("#"):
            cols, ready = {}, False
            continue
        if line.startswith("loop_"):
            cols, ready, loop = {}, False, 0
            continue
        if loop >= 0 and not line.startswith("_atom_site."):
            loop = -1
        if loop >= 0:
            f = _cif_field(line[11:] + " ")
            if f is not None:
                cols[f] = loop
            loop += 1
            ready = ("group_PDB" in cols and "label_atom_id" in cols and "label_comp_id" in cols and
                     ("auth_asym_id" in cols or "label_asym_id" in cols) and
                     ("auth_seq_id" in cols or "label_seq_id" in cols) and
                     "Cartn_x" in cols and "Cartn_y" in cols and "Cartn_z" in cols)
            continue
        if not ready:
            continue
        # tokens are read within the first 128 columns (maxlnwidth, FlexDataRead.h:67)
        spans = [(m.start(), m.end()) for m in re.finditer(r"[^ \t\r\n]+", line[:_CIF_LINE_WIDTH])]
        toks = [line[a:b] for a, b in spans]
        n = len(toks)
        if n < 1:
            continue
        g = cols["group_PDB"]
        if n <= g or not (toks[g].startswith("ATOM") or (hetatm and toks[g].startswith("HETATM"))):
            continue
        if "label_alt_id" in cols and (n <= cols["label_alt_id"] or toks[cols["label_alt_id"]][0] not in ".A"):
            continue
        a = cols["label_atom_id"]
        if n <= a or not (_cif_atom_is(toks[a], pa) or _cif_atom_is(toks[a], na)):
            continue
        need = [cols[k] for k in ("Cartn_x", "Cartn_y", "Cartn_z", "label_seq_id", "label_asym_id",
                                  "label_comp_id") if k in cols]
        if any(n <= c for c in need):
            continue
        ch = cols["auth_asym_id"] if "auth_asym_id" in cols and cols["auth_asym_id"] < n else cols["label_asym_id"]
        chain = toks[ch]
        chid = " " if chain[0] == "." else chain[0]
        insc = " "
        if "pdbx_PDB_ins_code" in cols and cols["pdbx_PDB_ins_code"] < n:
            insc = toks[cols["pdbx_PDB_ins_code"]][0]
        if insc == "?":
            insc = " "
        c = cols["label_comp_id"]
        comp = toks[c]
        resd = residue_code(line[spans[c][0]:spans[c][0] + 3].ljust(3))
        sq = cols["auth_seq_id"] if "auth_seq_id" in cols and cols["auth_seq_id"] < n else cols["label_seq_id"]
        resnum = _stoi(toks[sq])
        x, y, z = (_stof(toks[cols[k]]) for k in ("Cartn_x", "Cartn_y", "Cartn_z"))
        end = False
        if "pdbx_PDB_model_num" in cols and cols["pdbx_PDB_model_num"] < n:
            model = toks[cols["pdbx_PDB_model_num"]]
            if model != model_prev and xyz:
                end = True
            model_prev = model
        if chain != chain_prev:
            if xyz:
                end = True
            chord += 1
        if end:
            break
        if not xyz:
            chain0, model0 = chain, (model_prev or "")
        chain_prev = chain
        restype = 1 if (len(comp) < 2 or comp[0] == "D") else -1
        moltype += restype
        if (max_length is not None and 0 < max_length < len(xyz)) or _BUFFER_LIMIT < len(xyz):
            return None
        word = ((ord(insc) & 0xff) << 16) | ((ord(chid) & 0xff) << 8) | (chord & 0xff)
        xyz.append((x, y, z))
        res.append(resd)
        rnum.append(resnum)
        ico.append(word if restype > 0 else -word)
    if len(xyz) < 1 or len(xyz) < min_length:
        return None
    ico = np.array(ico, dtype=np.int32)
    nucleic = moltype > 0
    ico[0] = abs(int(ico[0])) if nucleic else -abs(int(ico[0]))
    return PdbStructure(np.array(xyz, dtype=np.float32), "".join(res),
                        np.array(rnum, dtype=np.int32), ico, chain0, model0, nucleic)


def read_structure(source, **kw):
    """read_cif for paths ending in .cif or .cif.gz (the reference's file-type
    rule, InputFilelist.cpp:39-42), read_pdb otherwise."""
    if isinstance(source, str) and "\n" not in source and (source.endswith(".cif") or source.endswith(".cif.gz")):
        return read_cif(source, **kw)
    return read_pdb(source, **kw)
