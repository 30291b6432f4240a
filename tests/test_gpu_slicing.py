"""Spectrum slicing (windows of k): the union of the windows equals the lowest modes of the
dense FP64 eigendecomposition, with no repeated and no skipped mode."""
import numpy as np
import pytest
import torch

from oracle import hessian as OH
from tests.helpers import small_globule

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,total", [(24, 60), (10, 40)])
def test_windows_cover_lowest_modes_without_repeats(cuda, k, total):
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.slicing import slice_spectrum
    print("case: slicing 1200 atoms k=%d total=%d" % (k, total), flush=True)
    X = small_globule(1200, seed=41)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    Q = rigid_body_basis(A.xyz)
    r = slice_spectrum(A, total, k=k, Q=Q, tol_rel=1e-9, maxiter=400, verbose=True)
    for wd in r.windows:
        print("window", wd, flush=True)
    w = np.linalg.eigvalsh(OH.hessian_dense(X[A.order], 8.0, 1.0))
    want = w[6:6 + total]
    assert len(r.windows) >= 2
    assert r.eigval.shape == (total,)
    assert np.max(np.abs(r.eigval - want)) <= 1e-8 * w[-1], np.max(np.abs(r.eigval - want))
    assert np.max(r.rel_res) <= 1e-9
    V = r.eigvec
    G = (V @ V.T).cpu().numpy()
    assert np.max(np.abs(G - np.eye(total))) <= 1e-8          # a repeated mode would show as ~1 off the diagonal
    rig = (V @ Q.T).abs().max().item()
    assert rig <= 1e-8
