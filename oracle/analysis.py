"""Mode analysis restated (oracle): square fluctuations and cross-correlation maps.

* ``sqfluct``            Burn/Visualisation/T2.py:43-55
* ``cross_correlation``  Burn/Visualisation/T2.py:58-116 (the hessian_CC loop; the plotting
                         branch :123-156 draws figures and is not restated) and the
                         log-modulus S_LogModulusS of Burn/Visualisation/T1.py:44-52

Modes are [k, N, 3] float64 (mode, atom, xyz). The reference forms each map in float64,
accumulates it into a float32 tensor and transforms it in float32; so does this restatement.
"""
import numpy as np


def sqfluct(eigval, eigvec):
    """(SqFluctPerMode [k, N], SqFluct [N], RatioSqFPM [k, N], RatioVariance [k]) (T2.py:43-55)."""
    lam = np.asarray(eigval, dtype=np.float64)
    V = np.asarray(eigvec, dtype=np.float64)
    rv = 1.0 / np.sqrt(np.abs(lam))
    rv = rv / rv.max()
    sfpm = np.sqrt(np.sum(V * V, axis=2))
    ratio = rv[:, None] * sfpm
    return sfpm, ratio.sum(axis=0), ratio, rv


def log_modulus(S, precision=1):
    """S_LogModulusS (T1.py:44-52) on a float32 array, in float32."""
    S = np.asarray(S, dtype=np.float32) * np.float32(100.0 ** precision)
    return (np.sign(S) * np.log2(np.abs(S) + np.float32(1.0))).astype(np.float32)


def pair_list(c_mode, self_only=True):
    """The pairs of T2.py:90-93: (i, i) for i < c_mode, or the sorted upper triangle."""
    if self_only:
        return sorted((i, i) for i in range(c_mode))
    return sorted((i, j) for i in range(c_mode) for j in range(i, c_mode))


def cross_correlation(eigval, eigvec, pairs, precision=2):
    """float32 [P, N, N]: log_modulus of float32((V_i V_j^T) lambda_i lambda_j) (T2.py:95-116)."""
    lam = np.asarray(eigval, dtype=np.float64)
    V = np.asarray(eigvec, dtype=np.float64)
    N = V.shape[1]
    out = np.zeros((len(pairs), N, N), dtype=np.float32)
    for p, (i, j) in enumerate(pairs):
        out[p] += ((V[i] @ V[j].T) * lam[i] * lam[j]).astype(np.float32)
    return log_modulus(out, precision)
