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
