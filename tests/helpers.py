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
