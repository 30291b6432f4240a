"""Small seeded structures and comparison helpers shared by the tests."""
import numpy as np


def small_globule(n, seed=0, density=0.058, dmin=2.2):
    """numpy dart throwing (test inputs only; independent of the library generator)."""
    rng = np.random.default_rng(seed)
    R = (3.0 * n / (4.0 * np.pi * density)) ** (1.0 / 3.0)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-R, R, size=3)
        if p @ p > R * R:
            continue
        if pts:
            P = np.asarray(pts)
            if np.min(np.sum((P - p) ** 2, axis=1)) < dmin * dmin:
                continue
        pts.append(p)
    return np.asarray(pts)


def subspace_sines(U, V):
    """Sines of the principal angles between span(U) and span(V) (orthonormal columns)."""
    s = np.linalg.svd(U.T @ V, compute_uv=False)
    s = np.clip(s, 0.0, 1.0)
    return np.sqrt(np.maximum(0.0, 1.0 - s ** 2))


def projected_system(n, i, kk, rng):
    """Orthonormal Q [n, i + 2] and a symmetric A whose restriction to span(Q[:, :i + 1]) is the
    thick-restart projected matrix T (Ritz values theta on rows < kk, arrowhead couplings
    betak to row kk, then tridiagonal alpha/beta) with A Q_i = sum_m T[m, i] Q_m + beta_i Q_{i+1};
    the orthogonal complement carries a random symmetric block."""
    Qf = np.linalg.qr(rng.standard_normal((n, n)))[0]
    m = i + 2
    alpha = rng.standard_normal(m)
    beta = rng.uniform(0.5, 2.0, m)
    theta = rng.standard_normal(kk)
    betak = rng.uniform(-1.0, 1.0, kk)
    T = np.zeros((m, m))
    for j in range(kk):
        T[j, j] = theta[j]
        T[j, kk] = T[kk, j] = betak[j]
    for j in range(kk, m):
        T[j, j] = alpha[j]
        if j + 1 < m:
            T[j, j + 1] = T[j + 1, j] = beta[j]
    R = rng.standard_normal((n - m, n - m))
    Af = np.zeros((n, n))
    Af[:m, :m] = T
    Af[m:, m:] = 0.5 * (R + R.T)
    A = Qf @ Af @ Qf.T
    return Qf[:, :m], A, alpha, beta, theta, betak, T
