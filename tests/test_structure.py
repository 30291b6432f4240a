"""Structure ingestion and result formats (CPU): reader rules on hand-written PDB/mmCIF
fixtures (first model, first alternate location, ATOM+HETATM, nm units), the driver's
centring, and the (k, N, 3) original-order eigenvector layout."""
import pickle

import numpy as np
import pytest

from inching_amd import structure as S

PDB = """\
HEADER    TEST
MODEL        1
ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N
ATOM      2  CA AALA A   1      11.639   6.071  -5.147  0.50  0.00           C
ATOM      3  CA BALA A   1      99.000  99.000  99.000  0.50  0.00           C
ATOM      4  C   ALA A   1      13.149   6.200  -5.100  1.00  0.00           C
HETATM    5  O   HOH A 101       1.000   2.000   3.000  1.00  0.00           O
TER
ENDMDL
MODEL        2
ATOM      1  N   ALA A   1      50.000  50.000  50.000  1.00  0.00           N
ENDMDL
END
"""

CIF = """\
data_test
#
loop_
_atom_site.group_PDB
_atom_site.id
_atom_site.type_symbol
_atom_site.label_atom_id
_atom_site.label_alt_id
_atom_site.label_comp_id
_atom_site.label_asym_id
_atom_site.label_seq_id
_atom_site.pdbx_PDB_ins_code
_atom_site.Cartn_x
_atom_site.Cartn_y
_atom_site.Cartn_z
_atom_site.pdbx_PDB_model_num
ATOM   1 N N   . ALA A 1 ? 11.104 6.134 -6.504 1
ATOM   2 C CA  A ALA A 1 ? 11.639 6.071 -5.147 1
ATOM   3 C CA  B ALA A 1 ? 99.000 99.000 99.000 1
ATOM   4 C C   . ALA A 1 ? 13.149 6.200 -5.100 1
HETATM 5 O O   . HOH B . ? 1.000 2.000 3.000 1
ATOM   6 N N   . ALA A 1 ? 50.000 50.000 50.000 2
#
"""

WANT_A = np.array([[11.104, 6.134, -6.504], [11.639, 6.071, -5.147], [13.149, 6.2, -5.1], [1.0, 2.0, 3.0]])


@pytest.mark.parametrize("name,text", [("t.pdb", PDB), ("t.cif", CIF)])
def test_reader_rules(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    xa = S.read_coordinates(str(p), unit="A")
    assert np.array_equal(xa, WANT_A)
    xn = S.read_coordinates(str(p))
    assert np.allclose(xn, WANT_A * 0.1, rtol=0, atol=1e-15)


def test_reader_errors(tmp_path):
    p = tmp_path / "empty.pdb"
    p.write_text("HEADER\nEND\n")
    with pytest.raises(S.StructureFormatError):
        S.read_coordinates(str(p))
    q = tmp_path / "x.xyz"
    q.write_text("1\n")
    with pytest.raises(S.StructureFormatError):
        S.read_coordinates(str(q))


def test_centre_rounds_mean_to_four_decimals():
    x = np.array([[0.12345, 1.0, 2.0], [0.0, 0.0, 0.0]])
    c = S.centre(x)
    assert np.array_equal(c, x - np.array([0.0617, 0.5, 1.0]))


def test_modes_in_atom_order_undoes_the_permutation(tmp_path):
    rng = np.random.default_rng(0)
    n, k = 50, 4
    V_orig = rng.standard_normal((k, n, 3))                 # modes in the original atom order
    order = rng.permutation(n).astype(np.int32)             # reordered position p holds atom order[p]
    undo = np.argsort(order)
    V_perm = V_orig[:, order, :].reshape(k, 3 * n)          # what the solver returns (rows)
    lam = np.arange(k) + 1.5
    h, vec = S.modes_in_atom_order(lam, V_perm, undo, plus_i=1.0)
    assert np.array_equal(h, lam - 1.0) and np.array_equal(vec, V_orig)
    h2, vec2 = S.modes_in_atom_order(lam, V_perm.T.copy(), undo, plus_i=1.0)   # reference column layout
    assert np.array_equal(vec2, V_orig)
    pv, pe = S.write_reference_pickles(str(tmp_path), "1abc", lam, V_perm, undo)
    with open(pe, "rb") as f:
        assert np.array_equal(pickle.load(f), V_orig)
    assert pv.endswith("Eigval_InchingTRLMHD_1abc_Linux_MI355X.pkl")
