"""Learned warm start: a coordinate-graph network that predicts a block spanning the
lowest non-trivial modes, used as the start of TRLM / JDM.

No counterpart in the reference: its solvers start from random vectors
(``Burn/ThickRestartLanczosHotellingDeflation/T1.py:117``, ``Burn/JacobiDavidsonHotellingDeflation/T1.py:462``).

Network (hidden 64, 4 graph layers, float16 weights and activations, float32
accumulation; every layer's forward and backward runs in the HIP kernels of
``inch_gnn_*``; torch is glue for the loss and the optimiser):

* per-atom input features from the centred coordinates scaled by the structure radius
  (position, radius, the 6 quadratic monomials, neighbour count);
* layer l: h <- relu(h W_self + mean_{j in N(i)} h_j W_nbr + b) over the ANM neighbour graph;
* head: 64 -> 3C (C vector fields per atom), i.e. a block P [3N, C] (C <= 32: the head emits 3C <= 96 channels).

Loss (subspace, so it is blind to the sign/rotation ambiguity of eigenvectors inside
near-degenerate clusters): with U the solver's lowest K non-trivial modes (orthonormal)
and P with the rigid-body block projected out, L = 1 - tr(U^T P (P^T P)^-1 P^T U) / K.

Use: Rayleigh-Ritz of the (FP64-orthonormalised) block with the true operator gives K
approximate eigenpairs; TRLM starts from a combination of them, JDM starts from the whole
block. The result is always the verified solve (float64 residual check); the cold start is
the fallback whenever the warm solve does not meet the tolerance.
"""
import json
import os

import numpy as np
import torch

from . import ops
from ._lib import call

HIDDEN = 64
LAYERS = 4
NFEAT = 16
F16, F32 = 2, 0
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_gnn.npz")

P_ = ops.P
SP = ops.stream_ptr


def features(xyz, rowptr, seg_atoms=None):
    """[N, 16] float32 input features of centred coordinates (any length unit). With
    ``seg_atoms`` (offsets [S+1] of contiguous atom ranges, e.g. the components of an
    assembly) every range is centred, scaled and degree-normalised on its own, as if it were
    a structure by itself."""
    x = xyz.to(torch.float64)
    deg = (rowptr[1:] - rowptr[:-1]).to(torch.float32)[:, None]
    if seg_atoms is None:
        c = x - x.mean(0, keepdim=True)
        R = torch.sqrt((c * c).sum(1).max()).clamp_min(1e-12)
        deg = deg / deg.mean()
    else:
        sizes = torch.as_tensor(np.diff(np.asarray(seg_atoms, dtype=np.int64)), device=x.device)
        seg = torch.repeat_interleave(torch.arange(sizes.numel(), device=x.device), sizes)
        cnt = sizes.to(torch.float64)[:, None]
        mean = torch.zeros((sizes.numel(), 3), dtype=torch.float64, device=x.device).index_add_(0, seg, x) / cnt
        c = x - mean[seg]
        r2 = (c * c).sum(1)
        R = torch.zeros(sizes.numel(), dtype=torch.float64, device=x.device).scatter_reduce_(
            0, seg, r2, reduce="amax", include_self=False).sqrt().clamp_min(1e-12)[seg][:, None]
        dm = torch.zeros((sizes.numel(), 1), dtype=torch.float32, device=x.device).index_add_(0, seg, deg)
        deg = deg / (dm / cnt.to(torch.float32))[seg]
    u = (c / R).to(torch.float32)
    r = torch.sqrt((u * u).sum(1, keepdim=True))
    q = torch.stack([u[:, 0] * u[:, 0], u[:, 1] * u[:, 1], u[:, 2] * u[:, 2],
                     u[:, 0] * u[:, 1], u[:, 1] * u[:, 2], u[:, 0] * u[:, 2]], 1)
    f = torch.cat([u, r, q, deg, torch.ones_like(r)], 1)          # 12
    f = torch.cat([f, torch.zeros((f.shape[0], NFEAT - f.shape[1]), device=f.device)], 1)
    return f.contiguous()


class _Graph:
    def __init__(self, rowptr, col):
        self.rowptr = rowptr
        self.col = col
        self.N = rowptr.numel() - 1


def _dense_fwd(X, W, b, Y32):
    N, Din = X.shape
    Dout = W.shape[1]
    call("inch_gnn_dense_fwd", N, Din, Dout, P_(X), P_(W), P_(b), 0, None, None, P_(Y32), SP())


def _dense_bwd(X, W, dY, dX, dW, db):
    N, Din = X.shape
    Dout = W.shape[1]
    from ._lib import lib
    work = torch.empty(int(lib().inch_gnn_work_size(Din, Dout)), dtype=torch.float32, device=X.device)
    call("inch_gnn_dense_bwd", N, Din, Dout, P_(X), P_(W), P_(dY), P_(dX), P_(dW), P_(db), P_(work), SP())


def _aggregate(g, X, xdt, Y, ydt, mode):
    call("inch_gnn_aggregate", g.N, X.shape[1], P_(g.rowptr), P_(g.col), P_(X), xdt, P_(Y), ydt, mode, SP())


class GraphLayer(torch.autograd.Function):
    """h_out = relu(h W_s + agg(h) W_n + b). Activations enter the kernels as float16,
    weights as float16, accumulation and gradients are float32."""

    @staticmethod
    def forward(ctx, h, Ws, Wn, b, g):
        N = h.shape[0]
        Dout = Ws.shape[1]
        h16 = h.to(torch.float16).contiguous()
        Ws16, Wn16 = Ws.to(torch.float16).contiguous(), Wn.to(torch.float16).contiguous()
        agg = torch.empty_like(h16)
        _aggregate(g, h16, F16, agg, F16, 0)
        a = torch.empty((N, Dout), dtype=torch.float32, device=h.device)
        c = torch.empty_like(a)
        _dense_fwd(h16, Ws16, b.contiguous(), a)
        _dense_fwd(agg, Wn16, None, c)
        pre = a + c
        ctx.save_for_backward(h16, agg, Ws16, Wn16, pre)
        ctx.g = g
        return torch.relu(pre)

    @staticmethod
    def backward(ctx, dout):
        h16, agg, Ws16, Wn16, pre = ctx.saved_tensors
        g = ctx.g
        dout = dout.to(torch.float32).contiguous()
        dpre = torch.empty_like(pre)
        call("inch_gnn_relu_bwd", pre.numel(), P_(pre), P_(dout), P_(dpre), SP())
        Din, Dout = Ws16.shape
        dh = torch.empty((h16.shape[0], Din), dtype=torch.float32, device=h16.device)
        dagg = torch.empty_like(dh)
        dWs = torch.empty((Din, Dout), dtype=torch.float32, device=h16.device)
        dWn = torch.empty_like(dWs)
        db = torch.empty(Dout, dtype=torch.float32, device=h16.device)
        db2 = torch.empty_like(db)
        _dense_bwd(h16, Ws16, dpre, dh, dWs, db)
        _dense_bwd(agg, Wn16, dpre, dagg, dWn, db2)
        dh2 = torch.empty_like(dh)
        _aggregate(g, dagg, F32, dh2, F32, 1)
        return dh + dh2, dWs, dWn, db, None


class Head(torch.autograd.Function):
    """y = h W + b (float16 operands, float32 out)."""

    @staticmethod
    def forward(ctx, h, W, b):
        h16 = h.to(torch.float16).contiguous()
        W16 = W.to(torch.float16).contiguous()
        y = torch.empty((h.shape[0], W.shape[1]), dtype=torch.float32, device=h.device)
        _dense_fwd(h16, W16, b.contiguous(), y)
        ctx.save_for_backward(h16, W16)
        return y

    @staticmethod
    def backward(ctx, dy):
        h16, W16 = ctx.saved_tensors
        Din, Dout = W16.shape
        dh = torch.empty((h16.shape[0], Din), dtype=torch.float32, device=h16.device)
        dW = torch.empty((Din, Dout), dtype=torch.float32, device=h16.device)
        db = torch.empty(Dout, dtype=torch.float32, device=h16.device)
        _dense_bwd(h16, W16, dy.contiguous().to(torch.float32), dh, dW, db)
        return dh, dW, db


class WarmStartNet:
    def __init__(self, channels=32, seed=0, device=None):
        self.C = int(channels)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        g = torch.Generator(device="cpu")
        g.manual_seed(int(seed))
        self.params = []
        din = NFEAT
        for _ in range(LAYERS):
            s = (2.0 / din) ** 0.5
            Ws = (torch.randn((din, HIDDEN), generator=g) * s).to(dev).requires_grad_()
            Wn = (torch.randn((din, HIDDEN), generator=g) * s).to(dev).requires_grad_()
            b = torch.zeros(HIDDEN, device=dev).requires_grad_()
            self.params += [Ws, Wn, b]
            din = HIDDEN
        Wo = (torch.randn((HIDDEN, 3 * self.C), generator=g) * (1.0 / HIDDEN) ** 0.5).to(dev).requires_grad_()
        bo = torch.zeros(3 * self.C, device=dev).requires_grad_()
        self.params += [Wo, bo]

    def forward(self, xyz, rowptr, col, seg_atoms=None):
        """Block P [3N, C] (float32; C <= 32: the head emits 3C channels) for coordinates in
        the matrix's atom order; ``seg_atoms``: disconnected parts featurised one by one (the
        graph layers never cross between them)."""
        gph = _Graph(rowptr, col)
        h = features(xyz, rowptr, seg_atoms)
        for l in range(LAYERS):
            Ws, Wn, b = self.params[3 * l:3 * l + 3]
            h = GraphLayer.apply(h, Ws, Wn, b, gph)
        y = Head.apply(h, self.params[-2], self.params[-1])        # [N, 3C]
        N = y.shape[0]
        return y.view(N, self.C, 3).permute(0, 2, 1).reshape(3 * N, self.C)

    def state(self):
        return {"p%d" % i: p.detach().cpu().numpy() for i, p in enumerate(self.params)} | {"C": np.array(self.C)}

    def save(self, path=WEIGHTS, meta=None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        st = self.state()
        if meta is not None:
            st["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(path, **st)

    @classmethod
    def load(cls, path=WEIGHTS, device=None):
        z = np.load(path)
        net = cls(channels=int(z["C"]), device=device)
        with torch.no_grad():
            for i, p in enumerate(net.params):
                p.copy_(torch.as_tensor(z["p%d" % i]))
        return net


def subspace_loss(P, U, Q=None):
    """1 - tr(U^T P (P^T P)^-1 P^T U) / K  (U [3N, K] orthonormal, P [3N, C] (C <= 32: the head emits 3C <= 96 channels).."""
    P = P.to(torch.float64)
    if Q is not None:
        P = P - Q.T @ (Q @ P)
    G = P.T @ P
    G = G + 1e-10 * torch.trace(G) / G.shape[0] * torch.eye(G.shape[0], dtype=G.dtype, device=G.device)
    M = U.T.to(torch.float64) @ P
    cap = torch.trace(M @ torch.linalg.solve(G, M.T))
    return 1.0 - cap / U.shape[1]


def train(net, samples, epochs=20, lr=3e-3, log=None):
    """samples: list of dicts with xyz, rowptr, col (device), U [3N, K] float64, Q [6, 3N]."""
    opt = torch.optim.Adam(net.params, lr=lr)
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for s in samples:
            opt.zero_grad()
            P = net.forward(s["xyz"], s["rowptr"], s["col"])
            L = subspace_loss(P, s["U"], s.get("Q"))
            L.backward()
            opt.step()
            tot += float(L.item())
        hist.append(tot / len(samples))
        if log:
            log("epoch %d  mean subspace loss %.4f" % (ep, hist[-1]))
    return hist


def warm_block(net, A, Q=None, lam_max=None):
    """Rayleigh-Ritz (float64, true operator) of the predicted block: returns
    (theta [C] ascending, X rows [C, n] float64, relative residuals)."""
    from .polish import polish
    with torch.no_grad():
        P = net.forward(A.xyz, A.rowptr, A.col)
    rows = P.T.to(torch.float64).contiguous()
    theta, X, info = polish(A, rows, Q=Q, lam_max=lam_max)
    return theta, X, info


def warm_start_vector(net, A, seg_atoms=None, seed=0):
    """A start vector [3N] (float64) for a lockstep solve: a combination of the predicted
    block's vector fields with positive random weights, featurised per component."""
    with torch.no_grad():
        P = net.forward(A.xyz, A.rowptr, A.col, seg_atoms)
    g = torch.Generator(device=P.device)
    g.manual_seed(int(seed) + 7)
    w = torch.rand(P.shape[1], generator=g, dtype=torch.float32, device=P.device) + 0.5
    return (P @ w).to(torch.float64)

