"""Per-component input features of the warm-start network (CPU, torch): featurising an
assembly's components together equals featurising each component alone."""
import numpy as np
import torch

from oracle import geometry as OG
from tests.helpers import small_globule


def _csr(X):
    ip, ix = OG.neighbour_csr(X, 8.0)
    return torch.as_tensor(ip), torch.as_tensor(ix)


def test_segment_features_equal_per_structure():
    from inching_amd.warmstart import features
    parts = [small_globule(300, seed=1), small_globule(450, seed=2) + 500.0, small_globule(200, seed=3) - 800.0]
    X = np.concatenate(parts)
    rp, _ = _csr(X)
    seg = np.cumsum([0] + [p.shape[0] for p in parts])
    together = features(torch.as_tensor(X), rp, seg)
    alone = torch.cat([features(torch.as_tensor(p), _csr(p)[0]) for p in parts])
    assert together.shape == (X.shape[0], 16)
    assert torch.allclose(together, alone, rtol=1e-5, atol=1e-6)
    # one segment over everything is the plain featurisation
    assert torch.allclose(features(torch.as_tensor(X), rp, [0, X.shape[0]]), features(torch.as_tensor(X), rp),
                          rtol=1e-6, atol=1e-7)
