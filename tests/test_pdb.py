"""PDB reader (gtalign_alpha_amd/pdb.py) against known answers fixed by the PDB
record format and by the reference reader's default options
(FlexDataRead.cpp:282-547, CLOptions.cpp:51-59, 91)."""
import gzip

import numpy as np
import pytest

from gtalign_alpha_amd import pdb


def atom(serial, name, resn, chain, resnum, x, y, z, alt=" ", ins=" ", rec="ATOM  "):
    # columns: 1-6 record, 7-11 serial, 13-16 name, 17 altLoc, 18-20 resName,
    # 22 chainID, 23-26 resSeq, 27 iCode, 31-54 x y z, 55-66 occupancy, B-factor
    return "%-6s%5d %-4s%1s%3s %1s%4d%1s   %8.3f%8.3f%8.3f%6.2f%6.2f" % (
        rec, serial, name, alt, resn, chain, resnum, ins, x, y, z, 1.0, 0.0)


def protein_chain(chain, n, start=1, resn="ALA", first_serial=1):
    lines = []
    for k in range(n):
        r = start + k
        lines.append(atom(first_serial + 2 * k, " N  ", resn, chain, r, k, 0.0, 0.0))
        lines.append(atom(first_serial + 2 * k + 1, " CA ", resn, chain, r, k + 0.5, 1.25, -2.0))
    return lines


def test_residue_codes():
    assert pdb.residue_code("MSE") == "M"
    assert pdb.residue_code("DAL") == "A"
    assert pdb.residue_code("SEC") == "U"
    assert pdb.residue_code("UNK") == "X"
    assert pdb.residue_code("  U") == "U"      # RNA name in column 20
    assert pdb.residue_code(" A ") == "A"      # one letter in column 19
    assert pdb.residue_code(" DA") == "X"      # DNA names map to X
    assert pdb.residue_code("A  ") == "A"
    assert pdb.residue_code("   ") == "X"


def test_first_chain_ca_atoms_known_answer():
    lines = ["HEADER    TEST", "MODEL        7"]
    lines += protein_chain("A", 20)
    # alternate locations: 'A' kept, 'B' dropped
    lines.append(atom(100, " CA ", "GLY", "A", 21, 9.0, 8.0, 7.0, alt="A", ins="B"))
    lines.append(atom(101, " CA ", "GLY", "A", 21, 1.0, 1.0, 1.0, alt="B", ins="B"))
    # HETATM skipped by default
    lines.append(atom(102, " CA ", "MSE", "A", 22, 3.0, 3.0, 3.0, rec="HETATM"))
    lines.append(atom(103, " CA ", "TRP", "A", -3, -123.456, 45.5, 0.001))
    lines.append("TER     104      TRP A  -3")
    lines += protein_chain("B", 30, first_serial=200)
    lines.append("END")
    s = pdb.read_pdb("\n".join(lines) + "\n")
    assert s is not None and len(s) == 22
    assert s.chain == "A" and s.model == "7" and not s.nucleic
    assert s.residues == "A" * 20 + "GW"
    assert s.resnum.tolist() == list(range(1, 21)) + [21, -3]
    assert np.array_equal(s.coords[0], np.float32([0.5, 1.25, -2.0]))
    assert np.array_equal(s.coords[20], np.float32([9.0, 8.0, 7.0]))
    assert np.array_equal(s.coords[21], np.float32([-123.456, 45.5, 0.001]))
    # (ins << 16) | (chain << 8) | order, negated for protein residues
    assert s.ins_ch_ord[0] == -((ord(" ") << 16) | (ord("A") << 8) | 1)
    assert s.ins_ch_ord[20] == -(4342017)   # ins 'B'=66, chain 'A'=65, order 1
    # HETATM included on request, as its residue's code
    s2 = pdb.read_pdb("\n".join(lines), hetatm=True)
    assert s2.residues == "A" * 20 + "GMW"


def test_chain_change_and_model_end_close_structure():
    text = "\n".join(protein_chain("A", 25) + protein_chain("B", 40, first_serial=100))
    s = pdb.read_pdb(text)
    assert len(s) == 25 and s.chain == "A"
    text = "\n".join(["MODEL        1"] + protein_chain("A", 21) + ["ENDMDL", "MODEL        2"] +
                     protein_chain("A", 50, first_serial=100) + ["ENDMDL"])
    s = pdb.read_pdb(text)
    assert len(s) == 21 and s.model == "1"


def test_short_empty_and_length_limit():
    assert pdb.read_pdb("\n".join(protein_chain("A", 19))) is None
    assert len(pdb.read_pdb("\n".join(protein_chain("A", 19)), min_length=3)) == 19
    assert pdb.read_pdb("HEADER    NOTHING\nEND\n") is None
    with pytest.raises(ValueError):
        bad = atom(1, " CA ", "ALA", "A", 1, 0, 0, 0)
        pdb.read_pdb(bad[:30] + "     abc" + bad[38:] + "\n")


def test_max_length_rule_of_the_reference_reader():
    """PMBatchStrData::AddOneResidue (PMBatchStrData.h:368-386) abandons the
    structure when a residue arrives while more than maxstrlen residues are
    stored: with --dev-max-length 4000 a 4,001-residue chain is kept and a
    4,002-residue one dropped; no limit for queries."""
    def ca_chain(n):
        return "\n".join(atom(k + 1, " CA ", "ALA", "A", k % 9999 + 1, k * 0.001, 0.0, 0.0) for k in range(n))
    kept = pdb.read_pdb(ca_chain(4001), max_length=pdb.DEV_MAX_LENGTH)
    assert kept is not None and len(kept) == 4001
    assert pdb.read_pdb(ca_chain(4002), max_length=pdb.DEV_MAX_LENGTH) is None
    assert len(pdb.read_pdb(ca_chain(4002))) == 4002
    assert pdb.read_pdb(ca_chain(21), max_length=19) is None
    assert len(pdb.read_pdb(ca_chain(20), max_length=19)) == 20


def test_nucleic_acid_c3_prime():
    lines = []
    for k in range(24):
        lines.append(atom(2 * k + 1, " P  ", "  G", "R", k + 1, 0, 0, 0))
        lines.append(atom(2 * k + 2, " C3'", "  G", "R", k + 1, 1.0 * k, 2.0, 3.0))
    s = pdb.read_pdb("\n".join(lines))
    assert s.nucleic and s.residues == "G" * 24
    word = (ord(" ") << 16) | (ord("R") << 8) | 1
    assert (s.ins_ch_ord == word).all()       # nucleic residues keep a positive word


def test_chunk_from_files(tmp_path):
    rng = np.random.default_rng(3)
    lens = [33, 80, 21, 57]
    paths, ref = [], []
    for i, n in enumerate(lens):
        xyz = np.round(rng.normal(0, 20, (n, 3)), 3)
        lines = [atom(k + 1, " CA ", "LEU", "A", k + 1, *xyz[k]) for k in range(n)]
        p = tmp_path / ("s%d.pdb" % i)
        data = ("\n".join(lines) + "\nEND\n").encode()
        if i == 1:
            p = tmp_path / "s1.pdb.gz"
            data = gzip.compress(data)
        p.write_bytes(data)
        paths.append(str(p))
        ref.append(xyz.astype(np.float32))
    paths.append(str(tmp_path / "short.pdb"))
    (tmp_path / "short.pdb").write_text("\n".join(protein_chain("A", 5)))
    c, kept = pdb.read_chunk(paths)
    assert kept == [0, 1, 2, 3]
    c40, kept40 = pdb.read_chunk(paths, max_length=40)   # keeps chains of <= 41 residues
    assert kept40 == [0, 2] and c40.len.tolist() == [33, 21]
    assert c.len.tolist() == sorted(lens, reverse=True)
    for k in range(c.nstrs):
        assert np.array_equal(c.coords(k), ref[c.order[k]])
        assert c.residues(k) == "L" * lens[c.order[k]]
    assert c.resnum[:80].tolist() == list(range(1, 81))
    assert (c.ins_ch_ord < 0).all()


# ---- PDBx/mmCIF (FlexDataRead.cpp:553-908) ---------------------------------------

CIF_HEADER = """data_TEST
#
loop_
_atom_site.group_PDB
_atom_site.id
_atom_site.type_symbol
_atom_site.label_atom_id
_atom_site.label_alt_id
_atom_site.label_comp_id
_atom_site.label_asym_id
_atom_site.label_entity_id
_atom_site.label_seq_id
_atom_site.pdbx_PDB_ins_code
_atom_site.Cartn_x
_atom_site.Cartn_y
_atom_site.Cartn_z
_atom_site.occupancy
_atom_site.B_iso_or_equiv
_atom_site.auth_seq_id
_atom_site.auth_asym_id
_atom_site.pdbx_PDB_model_num
"""


def cif_row(serial, atom, comp, chain, seq, x, y, z, alt=".", ins="?", rec="ATOM", auth_chain=None,
            auth_seq=None, model=1):
    return "%s %d C %s %s %s %s 1 %d %s %.3f %.3f %.3f 1.00 0.00 %d %s %d" % (
        rec, serial, atom, alt, comp, chain, seq, ins, x, y, z,
        seq if auth_seq is None else auth_seq, auth_chain or chain, model)


def cif_protein_chain(chain, n, start=1, comp="ALA", first_serial=1, auth_chain=None, model=1):
    rows = []
    for k in range(n):
        r = start + k
        rows.append(cif_row(first_serial + 2 * k, "N", comp, chain, r, k, 0.0, 0.0, auth_chain=auth_chain,
                            model=model))
        rows.append(cif_row(first_serial + 2 * k + 1, "CA", comp, chain, r, k + 0.5, 1.25, -2.0,
                            auth_chain=auth_chain, model=model))
    return rows


def test_cif_known_answer():
    rows = cif_protein_chain("A", 20, auth_chain="B")
    rows.append(cif_row(100, "CA", "GLY", "A", 21, 9.0, 8.0, 7.0, alt="A", ins="C", auth_chain="B",
                        auth_seq=121))
    rows.append(cif_row(101, "CA", "GLY", "A", 21, 1.0, 1.0, 1.0, alt="B", auth_chain="B"))
    rows.append(cif_row(102, "CA", "MSE", "A", 22, 3.0, 3.0, 3.0, rec="HETATM", auth_chain="B"))
    rows.append(cif_row(103, '"CA"', "TRP", "A", 23, -123.456, 45.5, 0.001, auth_chain="B"))
    rows += cif_protein_chain("C", 30, first_serial=200)          # chain change ends the structure
    text = CIF_HEADER + "\n".join(rows) + "\n#\n"
    s = pdb.read_cif(text)
    assert s is not None and len(s) == 22
    assert s.chain == "B" and s.model == "1" and not s.nucleic    # auth_asym_id before label_asym_id
    assert s.residues == "A" * 20 + "GW"
    assert s.resnum.tolist() == list(range(1, 21)) + [121, 23]   # auth_seq_id
    assert np.array_equal(s.coords[20], np.float32([9.0, 8.0, 7.0]))
    assert np.array_equal(s.coords[21], np.float32([-123.456, 45.5, 0.001]))
    assert s.ins_ch_ord[0] == -((ord(" ") << 16) | (ord("B") << 8) | 1)
    assert s.ins_ch_ord[20] == -((ord("C") << 16) | (ord("B") << 8) | 1)
    s2 = pdb.read_cif(text, hetatm=True)
    assert s2.residues == "A" * 20 + "GMW"


def test_cif_model_change_rna_and_short():
    rows = cif_protein_chain("A", 25) + cif_protein_chain("A", 25, first_serial=100, model=2)
    s = pdb.read_cif(CIF_HEADER + "\n".join(rows) + "\n")
    assert len(s) == 25                                            # model change ends it
    assert pdb.read_cif(CIF_HEADER + "\n".join(cif_protein_chain("A", 19)) + "\n") is None
    # RNA: 1-letter names (nucleic), C3' atoms, quoted
    rna = [cif_row(k + 1, '"C3\'"', "AUGC"[k % 4], "R", k + 1, 3.0 * k, 1.0, 0.5) for k in range(24)]
    s = pdb.read_cif(CIF_HEADER + "\n".join(rna) + "\n")
    assert s.nucleic and s.residues == "AUGC" * 6 and s.ins_ch_ord[0] > 0
    # DNA names start with D: nucleic, code X
    dna = [cif_row(k + 1, '"C3\'"', "DA", "D", k + 1, 3.0 * k, 1.0, 0.5) for k in range(24)]
    s = pdb.read_cif(CIF_HEADER + "\n".join(dna) + "\n")
    assert s.nucleic and s.residues == "X" * 24


def test_cif_equals_pdb_reader(tmp_path):
    """The same chains written as PDB and as mmCIF read to the same chunk."""
    rng = np.random.default_rng(11)
    names = ["ALA", "GLY", "TRP", "MSE", "LYS"]
    for n in (20, 57):
        xyz = np.round(rng.normal(scale=15.0, size=(n, 3)), 3)
        comps = [names[i % 5] for i in range(n)]
        plines = [atom(k + 1, " CA ", comps[k], "A", k + 5, *xyz[k]) for k in range(n)] + ["END"]
        crows = [cif_row(k + 1, "CA", comps[k], "A", k + 5, *xyz[k]) for k in range(n)]
        (tmp_path / ("s%d.pdb" % n)).write_text("\n".join(plines) + "\n")
        (tmp_path / ("s%d.cif" % n)).write_text(CIF_HEADER + "\n".join(crows) + "\n#\n")
    a, ka = pdb.read_chunk([str(tmp_path / "s20.pdb"), str(tmp_path / "s57.pdb")])
    b, kb = pdb.read_chunk([str(tmp_path / "s20.cif"), str(tmp_path / "s57.cif")])
    assert ka == kb == [0, 1]
    for f in ("len", "x", "y", "z", "rsd", "resnum", "ins_ch_ord"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
