"""Mode analysis on the device (inch_sqfluct, inch_cross_correlation) against the oracle
(oracle/analysis.py, a restatement of Burn/Visualisation/T2.py:43-116) and against the
closed forms that pin it, on modes this library computed and on dense-eigh modes."""
import numpy as np
import pytest
import torch

from oracle import analysis as OA
from oracle import hessian as OH
from tests.helpers import small_globule

pytestmark = pytest.mark.gpu


def _dense_modes(n=60, seed=3):
    X = small_globule(n, seed=seed)
    H = OH.hessian_dense(X, 8.0, 0.0)
    w, U = np.linalg.eigh(H)
    return H, w, U.T.reshape(-1, n, 3).copy()


def test_sqfluct_matches_oracle_and_pseudo_inverse(cuda):
    from inching_amd import analysis
    print("case: sqfluct 60 atoms, all non-rigid modes", flush=True)
    H, w, V = _dense_modes()
    nz = w > 1e-8
    sfpm, sqf, ratio, rv = analysis.square_fluctuations(torch.as_tensor(w[nz], device="cuda"),
                                                        torch.as_tensor(V[nz], device="cuda"))
    o = OA.sqfluct(w[nz], V[nz])
    for d, r in zip((sfpm, sqf, ratio, rv), o):
        assert np.allclose(d.cpu().numpy(), r, rtol=1e-13, atol=1e-15)
    pinv = np.linalg.pinv(H, rcond=1e-10)
    blk = np.array([np.trace(pinv[3 * a:3 * a + 3, 3 * a:3 * a + 3]) for a in range(60)])
    assert np.allclose((ratio.cpu().numpy() ** 2).sum(0), w[nz].min() * blk, rtol=1e-10)


@pytest.mark.parametrize("precision", [0, 2])
def test_cross_correlation_matches_oracle(cuda, precision):
    from inching_amd import analysis
    print("case: cross-correlation 300 atoms precision %d" % precision, flush=True)
    rng = np.random.default_rng(5)
    k, n = 7, 300
    V = rng.standard_normal((k, n, 3))
    V /= np.linalg.norm(V.reshape(k, -1), axis=1)[:, None, None]
    lam = np.sort(rng.uniform(0.01, 3.0, k))
    pairs = OA.pair_list(5, self_only=False)
    d = analysis.cross_correlation(lam, torch.as_tensor(V, device="cuda"), pairs, precision=precision).cpu().numpy()
    o = OA.cross_correlation(lam, V, pairs, precision=precision)
    assert d.dtype == np.float32 and d.shape == o.shape
    # same float32 value before the transform; log2 of the device and of numpy within 2 ulp
    assert np.max(np.abs(d - o) / np.maximum(np.abs(o), 1e-30)) <= 2.5e-7


def test_cross_correlation_h_squared_on_device(cuda):
    from inching_amd import analysis
    print("case: cross-correlation closed form", flush=True)
    H, w, V = _dense_modes()
    pairs = OA.pair_list(len(w), self_only=True)
    c = analysis.cross_correlation(w, torch.as_tensor(V, device="cuda"), pairs).double().cpu().numpy()
    inv = np.sign(c) * (2.0 ** np.abs(c) - 1.0) / 1e4
    H2 = H @ H
    ref = np.array([[np.trace(H2[3 * a:3 * a + 3, 3 * b:3 * b + 3]) for b in range(60)] for a in range(60)])
    assert np.max(np.abs(inv.sum(0) - ref)) <= 1e-6 * np.abs(ref).max()


def test_visualisation_shim_on_solver_modes(cuda):
    """The drop-in functions on modes from the solver, in atom order ([k, N, 3])."""
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes
    from inching_amd.structure import modes_in_atom_order
    from InchingLiteInteger.Burn.Visualisation import T2
    print("case: visualisation shim", flush=True)
    X = small_globule(700, seed=66)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    r = lowest_modes(A, 8, Q=rigid_body_basis(A.xyz), precision="mf64", tol_rel=1e-9)
    lam, modes_h = modes_in_atom_order(r.eigval, r.eigvec, A.undo, plus_i=1.0)
    lam_t, modes = torch.as_tensor(lam, device="cuda"), torch.as_tensor(modes_h, device="cuda")
    sfpm, sqf, ratio, rv = T2.Heigval_Heigvec_SqFluctPerMode_SqFluct_RatioSqFPM_RatioVar(lam_t, modes)
    assert sfpm.shape == (8, 700, 1) and ratio.shape == (8, 700, 1) and sqf.shape == (700,) and rv.shape == (8,)
    o = OA.sqfluct(lam, modes_h)
    assert np.allclose(sqf.cpu().numpy(), o[1], rtol=1e-13)
    tri, hcc = T2.Heigval_Heigvec_HccPairlist_HccBatch(lam_t, modes, plot_CC=False, c_mode=3, SelfOnly=False)
    assert torch.equal(tri, torch.triu_indices(3, 3).T)
    assert hcc.shape == (6, 700, 700) and hcc.dtype == torch.float32
    o = OA.cross_correlation(lam, modes_h, OA.pair_list(3, self_only=False))
    assert np.max(np.abs(hcc.cpu().numpy() - o) / np.maximum(np.abs(o), 1e-30)) <= 2.5e-7
    tri, hcc = T2.Heigval_Heigvec_HccPairlist_HccBatch(lam_t, modes, c_mode=4)
    assert hcc.shape == (4, 700, 700) and tri.shape == (10, 2)


def test_cross_correlation_rejects_bad_pairs(cuda):
    from inching_amd import analysis
    from inching_amd._lib import InchingError
    V = torch.zeros((3, 10, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(InchingError):
        analysis.cross_correlation(np.ones(3), V, [(0, 3)])
    with pytest.raises(InchingError):
        analysis.cross_correlation(np.ones(3), V, [(-1, 0)])
