"""Jacobi-Davidson with a deflated GMRES correction solve, on the device.

Follows ``Burn/JacobiDavidsonHotellingDeflation/T1.py:382-757`` (``S_HeigvalJDMHD_HeigvecJDMHD``)
and ``IterativeSolvers.py:216-305`` (``gmresDeflate``):

* search space V (at most jmax = factor*k vectors) with projected matrix G = V^T A V
  (host eigh each outer iteration, as the reference does);
* Ritz pair (theta, u) of the smallest Ritz value, residual r = A u - theta u; pairs with
  ||r|| <= tol are accepted, locked (Q) and deflated from V;
* restart to jmin = k vectors when V is full;
* correction: GMRES(20), at most ``corr_maxiter`` iterations, on (A - sigma I)(I - u u^T) with
  the residual measured through (I - u u^T), sigma = theta once ||r|| < 1e-3 else theta - gap;
* the correction is orthogonalised against the locked vectors (MGS) and V (ICGS, <= 3 passes).

The operator is A (+ shift Q^T Q for the Hotelling deflation), float64 stored or
matrix-free (``Operator`` with precision 'f64' or 'mf64'). The correction solve may run on
a second operator (``inner``): float32 stored values with float64 vectors, the mixed
precision form; it only shapes the search direction, so Ritz values, residuals and
acceptance stay on the float64 operator. Device vectors are rows of row-major blocks;
Gram products use the multi-dot kernel.
"""
import time

import numpy as np
import torch

from . import ops


class JDMResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class JDM:
    def __init__(self, op, k, factor=2, restart=20, inner=None):
        if factor <= 1:
            raise AssertionError("ABORTED. User_WorkspaceSizeFactor: int > 1")
        self.op = op
        self.inner = op if inner is None else inner
        if self.inner.n != op.n:
            raise ValueError("the correction operator has %d rows, the operator %d" % (self.inner.n, op.n))
        self.k = int(k)
        self.n = op.n
        self.jmax = self.k * int(factor)
        self.jmin = min(self.n, self.k)
        dev = op.A.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.V = torch.zeros((self.jmax + 1, self.n), **f64)
        self.Vtmp = torch.zeros((self.jmax + 1, self.n), **f64)
        self.Qa = torch.zeros((self.jmin, self.n), **f64)
        self.u = torch.zeros(self.n, **f64)
        self.r = torch.zeros(self.n, **f64)
        self.z = torch.zeros(self.n, **f64)
        self.Az = torch.zeros(self.n, **f64)
        self.restart = int(restart)
        self.Vg = torch.zeros((self.restart, self.n), **f64)
        self.w = torch.zeros(self.n, **f64)
        self.x = torch.zeros(self.n, **f64)
        self.mx = torch.zeros(self.n, **f64)
        self.h = torch.zeros(max(self.jmax + 1, self.restart + 1, self.jmin + 1), **f64)
        self.s1 = torch.zeros(1, **f64)
        self.Hd = torch.zeros((self.restart, self.restart + 1), **f64)   # row j = column j of H
        self.ws = ops.Workspace(self.h.numel(), dev)

    def _n_spmv(self):
        return self.op.n_spmv + (self.inner.n_spmv if self.inner is not self.op else 0)

    # -- small device helpers ------------------------------------------------
    def _dot(self, a, b):
        ops.multidot(a.view(1, -1), 1, b, self.s1, self.ws)
        return float(self.s1.item())

    def _nrm(self, a):
        ops.gemv_update(a.view(1, -1), 0, None, a, 0.0, nrm2_out=self.s1, ws=self.ws)
        return float(self.s1.item())

    def _combine(self, rows, m, W, out):
        ops.rr_combine(rows, m, torch.as_tensor(np.ascontiguousarray(W), device=rows.device), out)

    def _proj_apply(self, x, out, uu, sigma):
        """out = (A - sigma) (I - uu uu^T) x ; returns out (x is left projected)."""
        c = self._dot(uu, x)
        ops.axpby(-c, uu, 1.0, x)
        self.op.apply(x, out)
        ops.axpby(-sigma, x, 1.0, out)
        return out

    def _project(self, uu, x):
        """x -= (uu . x) uu with the dot kept on the device."""
        ops.multidot(uu.view(1, -1), 1, x, self.s1, self.ws)
        ops.axpy_dev(self.s1, -1.0, uu, x)

    def _gmres(self, b, uu, sigma, tol, maxiter, rng):
        """gmresDeflate: returns the projected iterate (the reference returns mx).

        The Arnoldi cycle stays on the device: projections use device dots, the Hessenberg
        column and the next basis vector come from the dot and norm kernels without a host
        round trip; the (R+1) x R Hessenberg matrix is read once per cycle for the
        least-squares solve, and the residual norm once per restart."""
        n, R = self.n, self.restart
        x, mx = self.x, self.mx
        x.copy_(torch.as_tensor(rng.random(n) * (tol / n), device=x.device))
        b_norm = self._nrm(b)
        if b_norm == 0:
            mx.copy_(b)
            return mx
        atol = tol * b_norm
        Hd = self.Hd
        iters = 0
        while True:
            mx.copy_(x)
            self._project(uu, mx)
            self.inner.apply(mx, self.w)
            ops.axpby(-sigma, mx, 1.0, self.w)
            self._project(uu, self.w)
            # w <- b - w
            ops.axpby(1.0, b, -1.0, self.w)
            r_norm = self._nrm(self.w)
            if r_norm <= atol or iters >= maxiter:
                break
            e = np.zeros(R + 1)
            ops.axpby(1.0 / r_norm, self.w, 0.0, self.Vg[0])
            e[0] = r_norm
            Hd.zero_()
            for j in range(R):
                z = self.z
                z.copy_(self.Vg[j])
                self._project(uu, z)
                self.inner.apply(z, self.Az)
                ops.axpby(-sigma, z, 1.0, self.Az)
                # the reference subtracts (uu . z) uu here, which is 0 after the projection
                ops.multidot(self.Vg, j + 1, self.Az, Hd[j, :j + 1], self.ws)
                ops.gemv_update(self.Vg, j + 1, Hd[j, :j + 1], self.Az, 1.0, nrm2_out=Hd[j, j + 1:j + 2], ws=self.ws)
                if j + 1 < R:
                    ops.scale_div(self.Az, Hd[j, j + 1:j + 2], self.Vg[j + 1])
            H = Hd.cpu().numpy().T                       # Hd holds the columns of H as rows
            # a zero subdiagonal entry (breakdown: the Krylov space is invariant) ends the
            # usable part of the cycle; columns after it were built from a division by zero
            m = R
            for j in range(R - 1):
                if not H[j + 1, j] != 0.0:
                    m = j + 1
                    break
            y = np.linalg.lstsq(H[:m + 1, :m], e[:m + 1], rcond=None)[0]
            ops.gemv_update(self.Vg, m, torch.as_tensor(y, device=x.device), x, -1.0)   # x += Vg^T y
            iters += R
        return mx

    def run(self, tol=1e-10, maxiter=1000, corr_maxiter=20, inter_tol=1e-3, gap=0.1, corr_factor=1e-4,
            seed=0, v0=None, V0=None, verbose=False):
        """V0 (optional): rows [m, n] orthonormal, a warm start search space (m <= jmax - 1)."""
        t0 = time.time()
        rng = np.random.default_rng(seed)
        k, n = self.k, self.n
        V, Qa = self.V, self.Qa
        eig = np.zeros(k)
        if v0 is None:
            v = torch.as_tensor(rng.random(n) - 0.5, device=V.device)
        else:
            v = torch.as_tensor(v0, dtype=torch.float64, device=V.device).reshape(-1).clone()
        v /= torch.linalg.norm(v)
        if V0 is not None:
            m0 = min(int(V0.shape[0]), self.jmax - 1)
            V[:m0].copy_(V0[:m0])
            Gd = torch.empty((m0, m0), dtype=torch.float64, device=V.device)
            AV = torch.empty(n, dtype=torch.float64, device=V.device)
            for j in range(m0):
                self.op.apply(V[j], AV)
                ops.multidot(V, m0, AV, self.h, self.ws)
                Gd[:, j] = self.h[:m0]
            G = Gd.cpu().numpy()
            G = 0.5 * (G + G.T)
        else:
            V[0].copy_(v)
            self.op.apply(V[0], self.Az)
            G = np.array([[self._dot(V[0], self.Az)]])
        nconv, jacc = 0, 0
        theta, cur_tol = None, None
        it = 0
        for it in range(maxiter):
            S, W = np.linalg.eigh(G, UPLO="L")
            while True:
                theta = S[0]
                m = W.shape[0]
                self._combine(V, m, W[:, :1], self.u.view(1, -1))
                self.op.apply(self.u, self.r)
                ops.axpby(-theta, self.u, 1.0, self.r)
                cur_tol = self._nrm(self.r)
                if verbose and it % 50 == 0:
                    print("JDM it %d  |r| %.3e  theta %.10f  converged %d" % (it, cur_tol, theta, nconv))
                sigma = theta - gap
                Qa[jacc].copy_(self.u)
                jcur = m
                if cur_tol > tol or (nconv != k - 1 and len(S) <= 1):
                    break
                eig[nconv] = theta
                nconv += 1
                jacc += 1
                self._combine(V, m, W[:, 1:], self.Vtmp)
                V[:W.shape[1] - 1].copy_(self.Vtmp[:W.shape[1] - 1])
                jcur = W.shape[1] - 1
                S = S[1:]
                G, W = np.diag(S), np.eye(S.shape[0])
                if nconv == k:
                    idx = np.argsort(eig)
                    torch.cuda.synchronize()
                    return JDMResult(eigval=eig[idx], eigvec=Qa[torch.as_tensor(idx, device=V.device)].clone(),
                                     iters=it, converged=nconv, n_spmv=self._n_spmv(), seconds=time.time() - t0)
            if S.shape[0] == self.jmax:
                self._combine(V, W.shape[0], W[:, :self.jmin], self.Vtmp)
                V[:self.jmin].copy_(self.Vtmp[:self.jmin])
                jcur = self.jmin
                S = S[:self.jmin]
                G, W = np.diag(S), np.eye(S.shape[0])
            shift = theta if cur_tol < inter_tol else sigma
            uu = self.u / torch.linalg.norm(self.u)
            ops.axpby(-1.0, self.r, 0.0, self.z)          # rhs = -r
            rhs = self.z.clone()
            z = self._gmres(rhs, uu, shift, cur_tol * corr_factor, corr_maxiter, rng).clone()
            for i in range(jacc + 1):                      # MGS against accepted + current
                c = self._dot(Qa[i], z)
                ops.axpby(-c, Qa[i], 1.0, z)
            r_pre = self._nrm(z)
            r1 = r_pre
            for _ in range(3):                             # ICGS against the search space
                ops.multidot(V, jcur, z, self.h, self.ws)
                ops.gemv_update(V, jcur, self.h, z, 1.0, nrm2_out=self.s1, ws=self.ws)
                r1 = float(self.s1.item())
                if r1 > r_pre / 2:
                    break
                r_pre = r1
            ops.axpby(0.0, z, 1.0 / r1, z)
            self.op.apply(z, self.Az)
            ops.multidot(V, jcur, self.Az, self.h, self.ws)
            col = self.h[:jcur].cpu().numpy()
            zz = self._dot(z, self.Az)
            G = np.vstack((np.hstack((G, col[:, None])), np.hstack((col[None, :], [[zz]]))))
            V[jcur].copy_(z)
        eig[nconv] = theta
        nconv += 1
        idx = np.argsort(eig)
        torch.cuda.synchronize()
        return JDMResult(eigval=eig[idx], eigvec=Qa[torch.as_tensor(idx[:self.jmin], device=V.device)].clone(),
                         iters=it, converged=nconv, n_spmv=self._n_spmv(), seconds=time.time() - t0)
