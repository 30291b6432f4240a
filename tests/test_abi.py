"""C ABI: the library loads, exports every declared symbol, and its host-only entry points
agree with the oracle (no GPU compute in this file)."""
import os
import re

import numpy as np
import pytest

from oracle import geometry as OG
from oracle import partition as OP
from oracle import rcm as OR
from tests.helpers import small_globule

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "inching.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(inch_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported(lib):
    from inching_amd import _lib
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), "missing export %s" % s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table and header disagree"
    assert lib.inch_abi_version() >= 1


def test_error_reporting(lib):
    import ctypes as C
    from inching_amd import _lib
    with pytest.raises(_lib.InchingError) as e:
        _lib.call("inch_synth_globule", -1, 10.0, 2.0, 0, 0, None)
    assert "bad args" in str(e.value)
    with pytest.raises(_lib.InchingError):
        out = np.empty((4000, 3))
        _lib.call("inch_synth_globule", 4000, 5.0, 2.2, 0, 100000, out.ctypes.data_as(C.c_void_p))


def test_rcm_host_matches_oracle(lib):
    from inching_amd import ops
    for seed, n in ((0, 500), (1, 1500)):
        X = small_globule(n, seed=seed)
        ip, ix = OG.neighbour_csr(X, 8.0)
        o_ref, u_ref = OR.rcm_order(ip, ix)
        o, u = ops.rcm_host(ip, ix)
        assert np.array_equal(o, o_ref) and np.array_equal(u, u_ref)
        # any seed order (here numpy's default argsort) is reproduced too
        so = np.argsort(np.diff(ip))
        o_ref2, _ = OR.rcm_order(ip, ix, seed_order=so)
        o2, _ = ops.rcm_host(ip, ix, seed_order=so)
        assert np.array_equal(o2, o_ref2)


def test_rcm_host_disconnected(lib):
    from inching_amd import ops
    rng = np.random.default_rng(5)
    X = np.concatenate([rng.uniform(0, 12, (150, 3)), rng.uniform(200, 212, (120, 3)), [[500.0, 0, 0]]])
    ip, ix = OG.neighbour_csr(X, 5.0)
    o_ref, _ = OR.rcm_order(ip, ix)
    o, _ = ops.rcm_host(ip, ix)
    assert np.array_equal(o, o_ref)


def test_batch_head_matches_oracle(lib):
    from inching_amd import ops
    for n in (1, 57, 100, 101, 98543, 300000):
        assert np.array_equal(ops.batch_head(n, 100), OP.batch_head(n, 100))


def test_synth_globule_deterministic_and_valid(lib):
    from inching_amd import synth
    a = synth.globule(3000, 11)
    b = synth.globule(3000, 11)
    c = synth.globule(3000, 12)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    R = synth.radius_for(3000)
    assert np.all(np.sum(a * a, 1) <= R * R + 1e-9)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(a).query(a, k=2)
    assert d[:, 1].min() >= 2.2 - 1e-12


def test_icosahedral_assembly_disconnected(lib):
    from inching_amd import synth
    G = synth.icosahedral_group()
    assert G.shape == (60, 3, 3)
    xyz, comp = synth.assembly(n_sub=60, atoms_per_sub=300, shell_radius=450.0)
    centres = np.stack([xyz[comp == c].mean(0) for c in range(60)])
    dc = np.linalg.norm(centres[:, None] - centres[None], axis=2) + np.eye(60) * 1e9
    assert dc.min() > 100.0
