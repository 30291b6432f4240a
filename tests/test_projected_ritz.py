"""The restart's projected eigenproblem on the device (``_projected_ritz_device``: padded to
the order where the batched eigh is fast) equals the host form (``_projected_ritz``) that
the in-cycle checks use, with and without the thick-restart arrowhead, for the largest and
the smallest pairs. Runs torch on the CPU."""
import numpy as np
import pytest
import torch

from inching_amd.batch import EIGH_MIN_ORDER, _projected_ritz, _projected_ritz_device


@pytest.mark.parametrize("S,m,k,filtered,restart", [(7, 40, 12, True, False), (7, 40, 12, False, True),
                                                     (5, 84, 24, True, True), (3, 30, 5, False, False),
                                                     (4, EIGH_MIN_ORDER, 9, True, True)])
def test_device_projected_ritz_equals_host(S, m, k, filtered, restart):
    rng = np.random.default_rng(S * 100 + m)
    alpha = rng.standard_normal((m + 3, S))
    beta = rng.random((m + 3, S)) + 0.1
    Sel_prev = np.sort(rng.standard_normal((S, k)), axis=1) if restart else None
    bk = 0.3 * rng.standard_normal((S, k)) if restart else None
    a, b = alpha[:m].copy(), beta[:m].copy()
    if restart:
        a[:k] = Sel_prev.T
        b[:k] = 0.0
    Sh, Wh, eh = _projected_ritz(a, b, bk, k, filtered)
    Sd, Wd, ed = _projected_ritz_device(torch.as_tensor(alpha), torch.as_tensor(beta), m, Sel_prev, bk, k,
                                        filtered)
    Wd = Wd.numpy()
    assert Wd.shape == (S, m, k) and Wd.flags["C_CONTIGUOUS"]
    np.testing.assert_allclose(Sd, Sh, rtol=0, atol=1e-12)
    sgn = np.sign(np.sum(Wh * Wd, axis=1, keepdims=True))     # eigenvectors up to sign
    np.testing.assert_allclose(Wd * sgn, Wh, rtol=0, atol=1e-10)
    np.testing.assert_allclose(ed * sgn[:, 0, :], eh, rtol=0, atol=1e-12)
    # the selected pairs are eigenpairs of the unpadded projected matrix
    T = np.zeros((S, m, m))
    i = np.arange(m)
    T[:, i, i] = a.T
    T[:, i[:-1], i[1:]] = b[:-1].T
    T[:, i[1:], i[:-1]] = b[:-1].T
    if restart:
        T[:, k, :k] = bk
        T[:, :k, k] = bk
    R = np.einsum("sij,sjq->siq", T, Wd) - Wd * Sd[:, None, :]
    assert np.abs(R).max() < 1e-11
