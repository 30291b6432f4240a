"""Chebyshev-Davidson subspace iteration (CDSI) on the device.

Follows ``Burn/ChebyshevDavidsonSubspaceIteration/T1.py:229-623``
(``S_HeigvalCDSIHD_HeigvecCDSIHD``, Zhou & Saad's single-vector Chebyshev-Davidson):

* the current Ritz vector x is filtered by the degree-m Chebyshev polynomial of
  (c - A) / e on [low, high] (``OOC5_KrylovAv_A_y_ynew_x_ChebyshevFilterSlim``, :196-226:
  y_1 = (c x - A x) / e, y_{k+1} = 2 (c y_k - A y_k) / e - y_{k-1}), which amplifies the
  spectrum below ``low``; every step is one fused operator launch;
* the filtered vector is orthogonalised against the search space (ICGS, at most three
  passes) and appended; its product with A extends Q = A V and the projected matrix G;
* Rayleigh-Ritz on the active (not yet converged) block, restart to
  max(nconv + 1, jmin + 5, min(jmin + nconv, jmax - 5)) vectors when the space is full;
* the smallest active Ritz pair is accepted when ||A x - theta x|| < tol * beta
  (beta = the largest |Ritz value| seen, as in the reference); ``low`` moves to the median
  of the active Ritz values (never below the spectrum's lower bound).

There is no Hotelling deflation (the reference asserts it away), so the rigid-body modes
(at PlusI) are among the lowest k returned.
"""
import time

import numpy as np
import torch

from . import ops
from .operator import Operator


class CDSIResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class CDSI:
    def __init__(self, A, k, factor=2, degree=80, precision="mf64"):
        if factor <= 1:
            raise AssertionError("ABORTED. We require User_WorkspaceSizeFactor > 1 ")
        self.A = A
        self.op = Operator(A, precision)
        self.n = A.shape[0]
        self.k = int(k)
        self.jmin = min(self.n, self.k)
        self.jmax = int(factor) * self.jmin
        self.degree = int(degree)
        dev = A.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.V = torch.zeros((self.jmax, self.n), **f64)
        self.Q = torch.zeros((self.jmax, self.n), **f64)
        self.Vt = torch.zeros((self.jmax, self.n), **f64)
        self.y = torch.empty(self.n, **f64)
        self.ynew = torch.empty(self.n, **f64)
        self.z = torch.empty(self.n, **f64)
        self.w = torch.empty(self.n, **f64)
        self.h = torch.zeros(self.jmax + 1, **f64)
        self.s1 = torch.zeros(1, **f64)
        self.ws = ops.Workspace(self.jmax + 1, dev)

    def _filter(self, x, low, high):
        """y = T_{m+1}((c - A) / e) x as the reference's slim filter (m = degree)."""
        e = (high - low) / 2.0
        c = (high + low) / 2.0
        z, y, ynew = self.z, self.y, self.ynew
        z.copy_(x)
        self.op.spmv(z, y, s1=-1.0 / e, c0=c)
        for _ in range(self.degree):
            self.op.spmv(y, ynew, s1=-2.0 / e, c0=c, z=z, c2=1.0)
            z, y, ynew = y, ynew, z
        self.z, self.y, self.ynew = z, y, ynew
        return y

    def _norm(self, v):
        ops.gemv_update(v.view(1, -1), 0, None, v, 0.0, nrm2_out=self.s1, ws=self.ws)
        return float(self.s1.item())

    def run(self, bound, tol=1e-10, maxiter=100000, seed=0, verbose=False):
        t0 = time.time()
        k, jmin, jmax = self.k, self.jmin, self.jmax
        V, Q = self.V, self.Q
        g = torch.Generator(device=V.device)
        g.manual_seed(int(seed))
        x = torch.rand(self.n, generator=g, dtype=torch.float64, device=V.device)
        x /= torch.linalg.norm(x)
        thb = float(bound[-1]) + 1e-5
        lower = float(bound[0])
        V[0].copy_(x)
        self.op.plain(V[0], Q[0])
        G = np.zeros((jmax, jmax))
        G[0, 0] = float(torch.dot(V[0], Q[0]))
        squeeze = (thb + G[0, 0]) / 2.0
        beta = max(1e-14, abs(G[0, 0]))
        tolr = tol * beta
        eig = np.zeros(jmin + 1)
        nconv, nconv1, jcur = 0, 1, 1
        it = 0
        for it in range(maxiter):
            v = self._filter(x, squeeze, thb)
            v.div_(self._norm(v))
            r_pre = 1.0
            for _ in range(3):                                   # ICGS against V[:jcur]
                ops.multidot(V, jcur, v, self.h, self.ws)
                ops.gemv_update(V, jcur, self.h, v, 1.0, nrm2_out=self.s1, ws=self.ws)
                r1 = float(self.s1.item())
                if r1 > r_pre / 2:
                    break
                r_pre = r1
            v.div_(r1)
            V[jcur].copy_(v)
            self.op.plain(V[jcur], Q[jcur])
            ops.multidot(V[nconv:], jcur + 1 - nconv, Q[jcur], self.h, self.ws)
            res = self.h[:jcur + 1 - nconv].cpu().numpy()
            G[jcur, nconv:jcur] = res[:-1]
            G[nconv:jcur + 1, jcur] = res
            G = 0.5 * (G + G.T)
            S, W = np.linalg.eigh(G[nconv:jcur + 1, nconv:jcur + 1], UPLO="L")
            if jcur + 1 >= jmax:
                j_restart = max(nconv1, jmin + 5, min(jmin + nconv, jmax - 5))
            else:
                j_restart = jcur + 1
            m_act = j_restart - nconv1 + 1
            Wd = torch.as_tensor(np.ascontiguousarray(W[:, :m_act]), device=V.device)
            m_old = jcur + 1 - nconv
            ops.rr_combine(V[nconv:], m_old, Wd, self.Vt)
            V[nconv:nconv + m_act].copy_(self.Vt[:m_act])
            ops.rr_combine(Q[nconv:], m_old, Wd, self.Vt)
            Q[nconv:nconv + m_act].copy_(self.Vt[:m_act])
            G[nconv:j_restart, nconv:j_restart] = np.diag(S[:m_act])
            G[nconv:, j_restart:] = 0.0
            G[j_restart:, nconv:] = 0.0
            beta1 = float(np.max(np.abs(S)))
            if beta1 > beta:
                beta = beta1
                tolr = tol * beta
            theta = S[0]
            x = V[nconv]
            ops.axpby(1.0, Q[nconv], 0.0, self.w)
            ops.axpby(-theta, x, 1.0, self.w)
            cur = self._norm(self.w)
            if verbose and it % 100 == 0:
                print("CDSI it %d |r| %.3e theta %.10f converged %d" % (it, cur, theta, nconv))
            if cur < tolr:
                eig[nconv] = theta
                nconv += 1
                nconv1 = nconv + 1
                if nconv >= jmin:
                    idx = np.argsort(eig[:jmin])
                    torch.cuda.synchronize()
                    return CDSIResult(eigval=eig[:jmin][idx], eigvec=V[torch.as_tensor(idx, device=V.device)].clone(),
                                      iters=it + 1, converged=nconv, n_spmv=self.op.n_spmv,
                                      seconds=time.time() - t0)
                x = V[nconv]
            jcur = j_restart
            med = float(np.median(S))
            squeeze = max(med, lower)
        idx = np.argsort(eig[:nconv])
        torch.cuda.synchronize()
        return CDSIResult(eigval=eig[:nconv][idx], eigvec=V[torch.as_tensor(idx, device=V.device)].clone(),
                          iters=it + 1, converged=nconv, n_spmv=self.op.n_spmv, seconds=time.time() - t0)
