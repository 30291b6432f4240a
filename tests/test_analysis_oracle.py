"""The mode-analysis oracle (oracle/analysis.py) pinned by closed forms on a small ANM
Hessian and by hand-built modes.

* With every non-rigid eigenpair of H, sum_m RatioSqFPM[m, a]^2 = lambda_min tr(H^+)_aa (the
  trace of atom a's 3 x 3 block of the pseudo-inverse): this fixes the 1 / sqrt|lambda|
  weighting, its normalisation by the largest weight and the per-atom magnitude axis.
* With every eigenpair, the self-pair maps undone by the inverse log-modulus and summed over
  the modes give the 3 x 3 block traces of H^2 = sum_m lambda_m^2 v_m v_m^T: this fixes the
  lambda_i lambda_j factor, the 100^precision scale, the log2 modulus and the contraction over
  xyz.
* Two hand-built modes living on single atoms fix the orientation of a cross pair (row =
  first mode's atom, column = second mode's atom) and the sign.
"""
import numpy as np

from oracle import analysis as OA
from oracle import hessian as OH
from tests.helpers import small_globule


def _eig(n=60, seed=3):
    X = small_globule(n, seed=seed)
    H = OH.hessian_dense(X, 8.0, 0.0)
    w, U = np.linalg.eigh(H)
    return H, w, U.T.reshape(-1, n, 3)


def test_sqfluct_pseudo_inverse_closed_form():
    H, w, V = _eig()
    nz = w > 1e-8
    assert nz.sum() == w.size - 6
    sfpm, sqf, ratio, rv = OA.sqfluct(w[nz], V[nz])
    pinv = np.linalg.pinv(H, rcond=1e-10)
    n = V.shape[1]
    blk = np.array([np.trace(pinv[3 * a:3 * a + 3, 3 * a:3 * a + 3]) for a in range(n)])
    assert np.allclose((ratio ** 2).sum(0), w[nz].min() * blk, rtol=1e-10, atol=1e-14)
    assert np.allclose((sfpm ** 2).sum(1), 1.0, atol=1e-12)           # unit modes
    assert np.allclose(sqf, ratio.sum(0)) and rv.max() == 1.0
    assert np.argmax(rv) == np.argmin(w[nz])


def test_cross_correlation_h_squared_closed_form():
    H, w, V = _eig()
    n = V.shape[1]
    pairs = OA.pair_list(len(w), self_only=True)
    cc = OA.cross_correlation(w, V, pairs, precision=2)
    assert cc.dtype == np.float32 and cc.shape == (len(w), n, n)
    c = cc.astype(np.float64)
    inv = np.sign(c) * (2.0 ** np.abs(c) - 1.0) / 1e4
    H2 = H @ H
    ref = np.array([[np.trace(H2[3 * a:3 * a + 3, 3 * b:3 * b + 3]) for b in range(n)] for a in range(n)])
    assert np.max(np.abs(inv.sum(0) - ref)) <= 1e-6 * np.abs(ref).max()


def test_cross_correlation_known_answer_orientation():
    # mode 0 lives on atom 1 (direction x), mode 1 on atom 3 (direction -x); eigenvalues 2, 0.5
    V = np.zeros((2, 5, 3))
    V[0, 1, 0] = 1.0
    V[1, 3, 0] = -1.0
    lam = np.array([2.0, 0.5])
    pairs = OA.pair_list(2, self_only=False)
    assert pairs == [(0, 0), (0, 1), (1, 1)]
    cc = OA.cross_correlation(lam, V, pairs, precision=1)
    # (0, 1): s[1, 3] = (1 * -1) * 2 * 0.5 * 100 = -100 -> -log2(101); everything else 0
    want = np.zeros((3, 5, 5), dtype=np.float32)
    want[0, 1, 1] = np.log2(np.float32(401.0))          # 1 * 4 * 100
    want[1, 1, 3] = -np.log2(np.float32(101.0))
    want[2, 3, 3] = np.log2(np.float32(26.0))           # 1 * 0.25 * 100
    assert np.array_equal(cc, want)
    assert OA.pair_list(3, self_only=True) == [(0, 0), (1, 1), (2, 2)]
