"""PDB reader: one structure per file, into the chunk format.

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
    """Chunk of the structures of the given PDB files (skipped files are left
    out); returns (chunk, indices of the sources that gave a structure).
    max_length: None for a query chunk (the reference reads queries with no
    limit), DEV_MAX_LENGTH for a reference chunk read as the reference does
    by default."""
    structs, kept = [], []
    for i, src in enumerate(sources):
        s = read_pdb(src, max_length=max_length, **kw)
        if s is not None:
            structs.append(s)
            kept.append(i)
    return chunk_from_structures(structs), kept
