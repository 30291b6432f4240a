"""Eigensolver restatements (oracle): spectrum bound, TRLM, IRLM, JDM.

Operators are passed as callables ``apply_A(v) -> A v`` (A = H + PlusI I, the
reference's matvec of ``Burn/Krylov/T3.py``); the optional rigid-body basis ``Q``
(m x n) enters as the Hotelling term ``+ shift Q^T Q v``
(``Burn/Orthogonalization/T3.py:62-164``). Start vectors come from a numpy
Generator passed in (the reference draws them from CuPy's global RNG, so iterates are
not reproducible; only converged eigenpairs are compared).

* ``spectrum_bound``  Burn/HermitianLanczos/T2.py:30-101
* ``trlm``            Burn/ThickRestartLanczosHotellingDeflation/T1.py:60-310 (driver),
                      :320-351 (tridiagonal/arrowhead selection), :415-522 (Lanczos steps)
* ``irlm``            Burn/ImplicitlyRestartedLanczosHotellingDeflation/T1.py:66-477
* ``jdm``             Burn/JacobiDavidsonHotellingDeflation/T1.py:338-757 and
                      IterativeSolvers.py:216-305 (gmresDeflate)
"""
import numpy as np

from .chebyshev import cheb_av


def spectrum_bound(apply_A, n, rng, n_tridiag=300):
    alpha = np.zeros(n_tridiag + 1)
    beta = np.zeros(n_tridiag)
    v = rng.random(n) - 0.5
    v /= np.linalg.norm(v)
    f = apply_A(v)
    a = v @ f
    f -= a * v
    alpha[0] = a
    v0 = np.zeros(n)
    for i in range(n_tridiag):
        b = np.linalg.norm(f)
        v0[:] = v
        v = f / b
        f = apply_A(v)
        f -= b * v0
        a = f @ v
        f -= a * v
        alpha[i + 1] += a
        beta[i] += b
    T = np.diag(alpha) + np.diag(beta, -1) + np.diag(beta, 1)
    w = np.linalg.eigvalsh(T)
    return float(np.min(w)), float(np.max(w))


def _select(alpha, beta, beta_k, k, filtered):
    t = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
    if beta_k is not None:
        t[k, :k] = beta_k
        t[:k, k] = beta_k
    S, W = np.linalg.eigh(t)
    idx = np.argsort(S)
    if not filtered:
        sel = idx[:k]                 # smallest k, ascending
    else:
        sel = idx[-k:][::-1]          # largest k, descending
    return S[sel], W[:, sel]


def trlm(apply_A, n, k, rng, maxiter=10000, tol=0.0, factor=2, Q=None, shift=10.0, pol=None, v0=None):
    """Thick-restart Lanczos with CGS2 (returns eigenvalues ascending, eigenvectors n x k, iters)."""
    ncv = min(max(factor * k, k + 32), n - 1)
    if tol == 0:
        tol = np.finfo(np.float64).eps
    if pol is None:
        op = apply_A
    else:
        op = lambda x: cheb_av(apply_A, x, pol)  # noqa: E731
    if Q is not None:
        base = op
        op = lambda x: base(x) + shift * (Q.T @ (Q @ x))  # noqa: E731
    alpha = np.zeros(ncv)
    beta = np.zeros(ncv)
    V = np.zeros((ncv, n))
    u = rng.random(n) if v0 is None else np.asarray(v0, dtype=np.float64).copy()
    V[0] = u / np.linalg.norm(u)

    def lanczos(i_start, i_end, u):
        v = V[i_start].copy()
        for i in range(i_start, i_end):
            u = op(v)
            alpha[i] = v @ u
            uu = V[:i + 1] @ u
            u -= V[:i + 1].T @ uu
            beta[i] = np.linalg.norm(u)
            uu = V[:i + 1] @ u
            u -= V[:i + 1].T @ uu
            if i >= i_end - 1:
                break
            v = u / beta[i]
            V[i + 1] = v
        return u

    u = lanczos(0, ncv, u)
    S, W = _select(alpha, beta, None, k, pol is not None)
    V[:k] = W.T @ V
    beta_k = beta[-1] * W[-1, :]
    res = np.linalg.norm(beta_k)
    it = 0
    for it in range(maxiter):
        beta[:k] = 0
        alpha[:k] = S
        u = u - V[:k].T @ (V[:k] @ u)
        u /= np.linalg.norm(u)
        V[k] = u
        u = op(V[k])
        alpha[k] = V[k] @ u
        u -= alpha[k] * V[k]
        u -= V[:k].T @ beta_k
        beta[k] = np.linalg.norm(u)
        V[k + 1] = u / beta[k]
        u = lanczos(k + 1, ncv, u)
        S, W = _select(alpha, beta, beta_k, k, pol is not None)
        V[:k] = W.T @ V
        beta_k = beta[-1] * W[-1, :]
        res = np.linalg.norm(beta_k)
        if res <= tol:
            break
    if pol is not None:
        S = np.array([V[i] @ apply_A(V[i]) for i in range(k)])
    idx = np.argsort(S)
    return S[idx], V[idx].T.copy(), it


def _orth_all_before(V, ix, before):
    s = V[:before] @ V[ix]
    V[ix] -= s @ V[:before]
    V[ix] /= np.linalg.norm(V[ix])


def irlm(apply_A, n, k, rng, maxiter=15000, tol=1e-8, Q=None, shift=10.0):
    """Implicitly restarted Lanczos (IRLM/T1.py:308-477); k >= 32 as asserted there."""
    assert k < n, "ABORTED. k must be smaller than n"
    assert k >= 32, "ABORTED. we did not test on less than 32 modes, as the number ritz vectors is too small."
    op = apply_A
    if Q is not None:
        op = lambda x: apply_A(x) + shift * (Q.T @ (Q @ x))  # noqa: E731
    p = min(max(2 * k, k + 32), n - 1)
    k = k + 2
    m = k + p
    assert n >= m
    V = np.zeros((m + 1, n))
    a = np.zeros(m)
    b = np.zeros(m)

    def lanczos(kk, pp):
        for j in range(pp):
            c = kk + j - 1
            if c == 0 or abs(b[kk + j - 2]) < tol:
                V[c] = rng.random(n)
                if c > 1:
                    _orth_all_before(V, c, c - 1)
                    _orth_all_before(V, c, c - 1)
                else:
                    V[c] /= np.sqrt(np.sum(V[c] * V[c]))
            u = op(V[c])
            if c > 0:
                h = V[c - 1:c + 1] @ u
                a[c] = h[1]
                b[c] = h[0]
                V[c + 1] = u - h @ V[c - 1:c + 1]
            else:
                h = np.sum(V[0] * u)
                a[0] = h
                V[1] = u - V[0] * h
            bb = np.sqrt(np.sum(V[c + 1] * V[c + 1]))
            if bb > tol:
                V[c + 1] /= bb
                b[c] = bb
            else:
                b[c] = 0
                continue
            _orth_all_before(V, c + 1, c)
            _orth_all_before(V, c + 1, c)

    def tridiag(al, be):
        return np.diag(al) + np.diag(be, 1) + np.diag(be, -1)

    def implicit_shift(a, b):
        T = tridiag(a, b[:k + p - 1])
        ev = np.linalg.eigvalsh(T)
        shifts = np.argsort(ev)[::-1][:p]
        Qacc = np.eye(k + p)
        for i in shifts:
            T = tridiag(a, b[:k + p - 1])
            lastb = b[k + p - 1:k + p].copy()
            Qj = np.linalg.qr(T - ev[i] * np.eye(k + p), mode="complete")[0]
            T = Qj.T @ (T @ Qj)
            a = np.diagonal(T).copy()
            b = np.concatenate((np.diagonal(T, 1), lastb))
            Qacc = Qacc @ Qj
        Qt = Qacc.T
        v = Qt[k] @ V[:k + p]
        V[:k] = Qt[:k] @ V[:k + p]
        V[k] = b[k - 1] * v + (b[k + p - 1] * Qt[k - 1, k + p - 1]) * V[k + p]
        bb = np.sqrt(np.sum(V[k] * V[k]))
        b[k - 1] = bb
        # A zero residual (an invariant subspace) leaves V[k] = 0 / 0 in the reference
        # (IRLM/T1.py:289-290); its next Lanczos call replaces V[k] by a random vector
        # because b[k - 1] < tol (:118-125), so skipping the normalisation here gives the
        # same iteration without the division by zero.
        if bb > 0.0:
            V[k] /= bb
            _orth_all_before(V, k, k)
            _orth_all_before(V, k, k)
        return a, b

    lanczos(1, k)
    it = 0
    for it in range(maxiter):
        lanczos(k + 1, p)
        a2, b2 = implicit_shift(a, b)
        a[:] = a2
        b[:] = b2
        if abs(b[k - 1]) < tol:
            b[k - 1] = 0.0
        if abs(b[k]) < tol:
            b[k] = 0.0
        if abs(b[k - 2]) < tol:
            break
    T = tridiag(a[:k], b[:k - 1])
    w, Z = np.linalg.eigh(T)
    V[:k - 1] = Z[:k - 1, :k - 1].T @ V[:k - 1]
    return w[:k - 2], V[:k - 2].T.copy(), it


def _gmres_deflate(matvec, b, uu, x0, tol, maxiter, restart=20):
    """gmresDeflate (IterativeSolvers.py:216-305): GMRES(restart) on (A - s I)(I - uu uu^T)."""
    n = b.shape[0]
    x = x0.copy()
    b_norm = np.linalg.norm(b)
    if b_norm == 0:
        return b
    atol = tol * b_norm
    restart = min(restart, n)
    iters = 0
    while True:
        mx = x.copy()
        mx -= (uu @ mx) * uu
        ax = matvec(mx)
        ax -= (uu @ ax) * uu
        r = b - ax
        r_norm = np.linalg.norm(r)
        if r_norm <= atol or iters >= maxiter:
            break
        Vg = np.zeros((n, restart))
        H = np.zeros((restart + 1, restart))
        e = np.zeros(restart + 1)
        Vg[:, 0] = r / r_norm
        e[0] = r_norm
        v = Vg[:, 0]
        for j in range(restart):
            z = v.copy()
            z -= (uu @ z) * uu
            u = matvec(z)
            u -= (uu @ z) * uu
            h = Vg[:, :j + 1].T @ u
            u -= Vg[:, :j + 1] @ h
            H[:j + 1, j] = h
            H[j + 1, j] = np.linalg.norm(u)
            if j + 1 < restart:
                v = u / H[j + 1, j]
                Vg[:, j + 1] = v
        y = np.linalg.lstsq(H, e, rcond=None)[0]
        x += Vg @ y
        iters += restart
    return mx


def jdm(apply_A, n, k, rng, tol=1e-10, maxiter=1000, corr_maxiter=20, factor=2, inter_tol=1e-3,
        gap=0.1, corr_factor=1e-4, Q=None, shift=10.0):
    """Jacobi-Davidson with deflated GMRES correction (JDM/T1.py:382-757)."""
    op = apply_A
    if Q is not None:
        op = lambda x: apply_A(x) + shift * (Q.T @ (Q @ x))  # noqa: E731
    jmax = k * factor
    jmin = min(n, k)
    eig = np.zeros(k)
    V = np.zeros((n, jmax + 1))
    Qa = np.zeros((n, jmin))
    v0 = rng.random((n, 1)) - 0.5
    v0 /= np.sqrt(np.sum(v0 * v0, axis=0))
    Av = op(v0[:, 0])
    V[:, 0] = v0[:, 0]
    G = v0.T @ Av[:, None]
    nconv = 0
    jacc = 0
    theta, cur_tol, u, r = None, None, None, None
    for it in range(maxiter):
        S, W = np.linalg.eigh(G, UPLO="L")
        while True:
            theta = S[0]
            u = V[:, :W.shape[0]] @ W[:, :1]
            r = op(u[:, 0])[:, None] - theta * u
            cur_tol = np.linalg.norm(r)
            sigma = theta - gap
            Qa[:, jacc] = u[:, 0]
            jcur = int(W.shape[0])
            if cur_tol > tol or (nconv != k - 1 and len(S) <= 1):
                break
            eig[nconv] = theta
            nconv += 1
            jacc += 1
            V[:, :W.shape[1] - 1] = V[:, :W.shape[0]] @ W[:, 1:]
            jcur = int(W.shape[1] - 1)
            S = S[1:]
            G, W = np.diag(S), np.eye(S.shape[0])
            if nconv == k:
                idx = np.argsort(eig)
                return eig[idx], Qa[:, idx], it
        if S.shape[0] == jmax:
            V[:, :jmin] = V[:, :W.shape[0]] @ W[:, :jmin]
            jcur = jmin
            S = S[:jmin]
            G, W = np.diag(S), np.eye(S.shape[0])
        s = theta if cur_tol < inter_tol else sigma
        matvec = lambda x, s=s: op(x) - s * x  # noqa: E731
        uu = u[:, 0] / np.linalg.norm(u[:, 0])
        ctol = cur_tol * corr_factor
        x0 = rng.random(n) * (ctol / n)
        z = _gmres_deflate(matvec, -r[:, 0], uu, x0, ctol, corr_maxiter)[:, None]
        for i in range(jacc + 1):                      # MGS against accepted + current
            z = z - (z.T @ Qa[:, i:i + 1]) * Qa[:, i:i + 1]
        Vc = V[:, :jcur]                               # ICGS against the search space
        r_pre = np.sqrt(abs(z.T @ z))
        r1 = r_pre
        for _ in range(3):
            z = z - Vc @ (Vc.T @ z)
            r1 = np.sqrt(abs(z.T @ z))
            if r1 > r_pre / 2:
                break
            r_pre = r1
        z = z / r1
        Az = op(z[:, 0])[:, None]
        G = np.vstack((np.hstack((G, Vc.T @ Az)), np.hstack((Az.T @ Vc, Az.T @ z))))
        V[:, jcur] = z[:, 0]
    eig[nconv] = theta
    nconv += 1
    idx = np.argsort(eig)
    return eig[idx], Qa[:, idx], maxiter
