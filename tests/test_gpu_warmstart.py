"""Learned warm start: device gradients against the CPU oracle, training reduces the loss,
and the trained network saves work against the cold start."""
import os

import numpy as np
import pytest
import torch

from oracle import geometry as OG
from oracle import gnn as ON
from tests.helpers import small_globule

pytestmark = pytest.mark.gpu


def _graph(n, seed):
    X = small_globule(n, seed=seed)
    ip, ix = OG.neighbour_csr(X, 8.0)
    return X, ip, ix


def test_layer_gradients_match_oracle(cuda):
    from inching_amd import warmstart as W
    print("case: gnn layer gradients", flush=True)
    X, ip, ix = _graph(300, 71)
    rng = np.random.default_rng(0)
    N, Din, Dout = 300, 16, 64
    h = rng.standard_normal((N, Din)).astype(np.float32)
    Ws = (rng.standard_normal((Din, Dout)) * 0.3).astype(np.float32)
    Wn = (rng.standard_normal((Din, Dout)) * 0.3).astype(np.float32)
    b = (rng.standard_normal(Dout) * 0.1).astype(np.float32)
    dout = rng.standard_normal((N, Dout))
    g = W._Graph(torch.as_tensor(ip, device="cuda"), torch.as_tensor(ix, device="cuda"))
    ht = torch.as_tensor(h, device="cuda").requires_grad_()
    Wst = torch.as_tensor(Ws, device="cuda").requires_grad_()
    Wnt = torch.as_tensor(Wn, device="cuda").requires_grad_()
    bt = torch.as_tensor(b, device="cuda").requires_grad_()
    out = W.GraphLayer.apply(ht, Wst, Wnt, bt, g)
    out.backward(torch.as_tensor(dout, dtype=torch.float32, device="cuda"))
    ref, cache = ON.layer_forward(ip, ix, h, Ws, Wn, b)
    dh, dWs, dWn, db = ON.layer_backward(ip, ix, cache, Ws, Wn, dout)
    # float16 operands, float32 accumulation: relative agreement ~1e-3 of the magnitudes
    assert np.allclose(out.detach().cpu().numpy(), ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())
    for got, want in ((ht.grad, dh), (Wst.grad, dWs), (Wnt.grad, dWn), (bt.grad, db)):
        got = got.cpu().numpy()
        assert np.linalg.norm(got - want) <= 5e-3 * np.linalg.norm(want)


def test_training_reduces_subspace_loss(cuda):
    from inching_amd import warmstart as W
    from inching_amd.batch import solve_batch
    from inching_amd.hessian import rigid_body_basis
    print("case: gnn short training", flush=True)
    structs = [small_globule(n, seed=80 + i) for i, n in enumerate([700, 900, 800])]
    r = solve_batch(structs, k=12, tol_rel=1e-8, maxiter=300)
    samples = []
    for s in range(3):
        a0, a1 = int(r.seg_atoms[s]), int(r.seg_atoms[s + 1])
        rows = r.A.rowptr[a0:a1 + 1]
        col = (r.A.col[int(rows[0]):int(rows[-1])] - a0).contiguous()
        xyz = r.A.xyz[a0:a1].contiguous()
        samples.append({"xyz": xyz, "rowptr": (rows - rows[0]).contiguous(), "col": col,
                        "U": r.eigvec[:, 3 * a0:3 * a1].T.contiguous(), "Q": rigid_body_basis(xyz)})
    net = W.WarmStartNet(channels=24, seed=0)
    hist = W.train(net, samples, epochs=25, lr=3e-3)
    assert hist[-1] < 0.8 * hist[0], hist


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(__file__)), "inching_amd",
                                                    "weights", "warmstart_gnn.npz")),
                    reason="trained weights not present")
def test_trained_network_saves_matvecs(cuda):
    """On an unseen globule the learned warm block, used as the initial search space of
    Jacobi-Davidson, reaches the same verified tolerance with fewer operator applications than
    the cold start, and both return the eigenvalues of the cold thick-restart Lanczos solve.
    (A single start vector built from the block gives TRLM no measurable gain: its restart
    count is set by the filter, not by the start.)"""
    from inching_amd import synth, warmstart as W
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes, lowest_modes_jdm, lowest_modes_jdm_warm
    print("case: warm vs cold on a 30k-atom globule", flush=True)
    net = W.WarmStartNet.load()
    X = synth.globule(30000, seed=991)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    Q = rigid_body_basis(A.xyz)
    k = 32
    ref = lowest_modes(A, k, Q=Q, precision="mf64", tol_rel=1e-8)
    cold = lowest_modes_jdm(A, k, Q=Q, precision="mf64", tol_rel=1e-8)
    warm = lowest_modes_jdm_warm(A, k, net, Q=Q, precision="mf64", tol_rel=1e-8)
    print("jdm cold %d matvecs, warm %d" % (cold.total_spmv, warm.total_spmv), flush=True)
    assert not warm.fallback
    assert np.max(warm.rel_res) <= 1e-8 and np.max(cold.rel_res) <= 1e-8
    assert np.allclose(warm.eigval, ref.eigval, rtol=0, atol=1e-9)
    assert np.allclose(cold.eigval, ref.eigval, rtol=0, atol=1e-9)
    assert warm.total_spmv < 0.9 * cold.total_spmv, (warm.total_spmv, cold.total_spmv)
