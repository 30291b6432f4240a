"""Structure ingestion and result formats around the eigensolver.

Input side (``InchingLiteInteger/util.py:90-108``, ``BasicPdbCifLoading``): the reference
reads PDB or mmCIF through openmm and takes every atom of the first model, with the first
alternate location of each atom, in nanometres; the driver then centres the coordinates as
``xyz -= around(mean(xyz), 4)`` (``commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:214-216``).
:func:`read_coordinates` is a dependency-free reader with those rules (ATOM and HETATM
records, first MODEL, first alternate location per atom).

Output side (``commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:396-407``): eigenvalues
are reported for H, i.e. ``eigval - PlusI``; eigenvectors as ``(k, N, 3)`` in the original
atom order, ``vec.T.reshape(k, N, 3)[:, undo, :]`` where ``undo`` is the inverse of the
3DRCM permutation. :func:`modes_in_atom_order` and :func:`write_reference_pickles` give
exactly those arrays and file names.
"""
import os
import pickle
import shlex

import numpy as np
import torch


class StructureFormatError(ValueError):
    pass


def _pdb_atoms(lines):
    seen = {}
    xyz = []
    in_model = False
    for ln in lines:
        rec = ln[:6]
        if rec.startswith("MODEL"):
            if in_model or xyz:
                break                      # second model: stop
            in_model = True
            continue
        if rec.startswith("ENDMDL"):
            if xyz:
                break
            continue
        if not (rec.startswith("ATOM") or rec.startswith("HETATM")):
            continue
        if len(ln) < 54:
            raise StructureFormatError("short coordinate record: %r" % ln.rstrip())
        name, alt = ln[12:16], ln[16]
        chain, resseq, icode = ln[21], ln[22:26], ln[26] if len(ln) > 26 else " "
        key = (chain, resseq, icode, ln[17:20], name)
        if alt != " ":
            if key in seen:
                continue                   # a later alternate location of an atom already read
            seen[key] = True
        try:
            xyz.append((float(ln[30:38]), float(ln[38:46]), float(ln[46:54])))
        except ValueError:
            raise StructureFormatError("bad coordinates: %r" % ln.rstrip())
    return np.asarray(xyz, dtype=np.float64).reshape(-1, 3)


def _cif_tokens(lines):
    for ln in lines:
        s = ln.strip()
        if not s or s.startswith("#"):
            continue
        if s.startswith(";"):
            raise StructureFormatError("multi-line text fields inside _atom_site are not supported")
        yield from shlex.split(s, posix=True)


def _cif_atoms(lines):
    cols = []
    toks = iter(_cif_tokens(lines))
    rows = []
    for t in toks:
        if t == "loop_":
            cols = []
            t = next(toks, None)
            while t is not None and t.startswith("_"):
                cols.append(t)
                t = next(toks, None)
            if not cols or not cols[0].startswith("_atom_site."):
                cols = []
                continue
            vals = []
            while t is not None and not t.startswith("_") and t != "loop_" and not t.startswith("data_"):
                vals.append(t)
                t = next(toks, None)
            if len(vals) % len(cols):
                raise StructureFormatError("_atom_site loop has a ragged row")
            rows = [vals[i:i + len(cols)] for i in range(0, len(vals), len(cols))]
            break
    if not rows:
        raise StructureFormatError("no _atom_site loop")
    c = {name[len("_atom_site."):]: i for i, name in enumerate(cols)}
    for need in ("Cartn_x", "Cartn_y", "Cartn_z"):
        if need not in c:
            raise StructureFormatError("_atom_site lacks %s" % need)

    def get(r, *names, default="."):
        for nm in names:
            if nm in c:
                return r[c[nm]]
        return default

    first_model = None
    seen = {}
    xyz = []
    for r in rows:
        grp = get(r, "group_PDB", default="ATOM")
        if grp not in ("ATOM", "HETATM"):
            continue
        model = get(r, "pdbx_PDB_model_num", default="1")
        if first_model is None:
            first_model = model
        elif model != first_model:
            break
        alt = get(r, "label_alt_id")
        key = (get(r, "auth_asym_id", "label_asym_id"), get(r, "auth_seq_id", "label_seq_id"),
               get(r, "pdbx_PDB_ins_code"), get(r, "auth_comp_id", "label_comp_id"),
               get(r, "auth_atom_id", "label_atom_id"))
        if alt not in (".", "?"):
            if key in seen:
                continue
            seen[key] = True
        try:
            xyz.append((float(r[c["Cartn_x"]]), float(r[c["Cartn_y"]]), float(r[c["Cartn_z"]])))
        except ValueError:
            raise StructureFormatError("bad coordinates in row %r" % (r,))
    return np.asarray(xyz, dtype=np.float64).reshape(-1, 3)


def read_coordinates(path, unit="nm"):
    """Coordinates [N, 3] float64 of a .pdb / .cif file (first model, first alternate
    location), in nanometres (``unit='nm'``, as the reference loader) or angstrom ('A')."""
    ext = os.path.splitext(path)[1].lower()
    with open(path, "r") as f:
        lines = f.read().splitlines()
    if ext == ".pdb" or ext == ".ent":
        xyz = _pdb_atoms(lines)
    elif ext in (".cif", ".mmcif"):
        xyz = _cif_atoms(lines)
    else:
        raise StructureFormatError("unknown structure format %r (use .pdb or .cif)" % ext)
    if xyz.shape[0] == 0:
        raise StructureFormatError("no atoms in %s" % path)
    if unit == "nm":
        return xyz * 0.1
    if unit in ("A", "angstrom"):
        return xyz
    raise ValueError("unit must be 'nm' or 'A'")


def centre(xyz):
    """The driver's centring: subtract the mean rounded to 4 decimals (keeps PDB digits exact)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    return xyz - np.around(xyz.mean(axis=0), decimals=4)


def modes_in_atom_order(eigval, eigvec, undo, plus_i=1.0):
    """(eigenvalues of H, eigenvectors (k, N, 3) in the original atom order).

    eigval: [k] eigenvalues of H + PlusI*I; eigvec: rows [k, 3N] (this package's layout) or
    columns [3N, k] (the reference's layout, recognised by shape) in the reordered frame;
    undo: inverse of the reordering permutation (``A.undo`` / the second output of
    ``X_KdCuthillMckeeOrder``)."""
    lam = eigval.detach().cpu().numpy() if torch.is_tensor(eigval) else np.asarray(eigval)
    V = eigvec.detach().cpu().numpy() if torch.is_tensor(eigvec) else np.asarray(eigvec)
    undo = np.asarray(undo, dtype=np.int64)
    n_atoms = undo.shape[0]
    k = lam.shape[0]
    if V.shape == (3 * n_atoms, k) and V.shape != (k, 3 * n_atoms):
        V = V.T
    if V.shape != (k, 3 * n_atoms):
        raise ValueError("eigvec shape %s does not match k=%d, N=%d" % (V.shape, k, n_atoms))
    return lam - plus_i, np.ascontiguousarray(V.reshape(k, n_atoms, 3)[:, undo, :])


def write_reference_pickles(folder, tag, eigval, eigvec, undo, plus_i=1.0, method="InchingTRLMHD",
                            platform="Linux", device="MI355X"):
    """Writes ``Eigval_<method>_<tag>_<platform>_<device>.pkl`` (eigenvalues of H) and
    ``Eigvec_...pkl`` ((k, N, 3), original atom order) as the reference's drivers do
    (pickle protocol 4). Returns the two paths."""
    lam, vec = modes_in_atom_order(eigval, eigvec, undo, plus_i)
    os.makedirs(folder, exist_ok=True)
    dev = device.replace(" ", "")
    p_val = os.path.join(folder, "Eigval_%s_%s_%s_%s.pkl" % (method, tag, platform, dev))
    p_vec = os.path.join(folder, "Eigvec_%s_%s_%s_%s.pkl" % (method, tag, platform, dev))
    with open(p_val, "wb") as f:
        pickle.dump(lam, f, protocol=4)
    with open(p_vec, "wb") as f:
        pickle.dump(vec, f, protocol=4)
    return p_val, p_vec
