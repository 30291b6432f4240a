"""Analysis of computed modes on the device (``Burn/Visualisation/T2.py:43-160``).

* ``square_fluctuations``  per-mode atom magnitudes |v_m[a]|, the relative variance of each
  mode (1 / sqrt|lambda|, normalised by its largest value), their product and its sum over
  the modes (``Heigval_Heigvec_SqFluctPerMode_SqFluct_RatioSqFPM_RatioVar``, :43-55).
* ``cross_correlation``    N x N maps, one per pair of modes (i, j): the log-modulus of
  lambda_i lambda_j (v_i[a] . v_j[b]) (the ``hessian_CC`` loop of
  ``Heigval_Heigvec_HccPairlist_HccBatch``, :58-116, with ``S_LogModulusS`` of
  ``Burn/Visualisation/T1.py:44-52``).

Modes are [k, N, 3] (mode, atom, xyz), float64 on the device; both are single kernels of
the library (``inch_sqfluct``, ``inch_cross_correlation``).
"""
import numpy as np
import torch

from . import ops
from ._lib import call

P = ops.P


def _modes(eigval, eigvec, device):
    ev = eigvec if isinstance(eigvec, torch.Tensor) else torch.as_tensor(np.asarray(eigvec))
    ev = ev.to(device=device, dtype=torch.float64).contiguous()
    if ev.dim() != 3 or ev.shape[2] != 3:
        raise ValueError("modes must be [k, N, 3]")
    lam = eigval if isinstance(eigval, torch.Tensor) else torch.as_tensor(np.asarray(eigval))
    lam = lam.to(device=device, dtype=torch.float64).reshape(-1).contiguous()
    if lam.numel() != ev.shape[0]:
        raise ValueError("eigval has %d entries for %d modes" % (lam.numel(), ev.shape[0]))
    return lam, ev


def square_fluctuations(eigval, eigvec, device=None):
    """(sfpm [k, N], sqfluct [N], ratio_sfpm [k, N], ratio_var [k]) of the modes, float64."""
    ops.require_gpu()
    device = device or (eigvec.device if isinstance(eigvec, torch.Tensor) and eigvec.is_cuda
                        else torch.device("cuda", torch.cuda.current_device()))
    lam, ev = _modes(eigval, eigvec, device)
    k, N = ev.shape[0], ev.shape[1]
    f64 = dict(dtype=torch.float64, device=device)
    ratio_var = torch.empty(k, **f64)
    sfpm = torch.empty((k, N), **f64)
    ratio = torch.empty((k, N), **f64)
    sqf = torch.empty(N, **f64)
    call("inch_sqfluct", k, N, P(lam), P(ev), P(ratio_var), P(sfpm), P(ratio), P(sqf), ops.stream_ptr())
    return sfpm, sqf, ratio, ratio_var


def mode_pairs(c_mode, self_only=True):
    """The pair list of the reference: (i, i) for i < c_mode, or every (i, j), i <= j, in
    row-major upper-triangle order."""
    if self_only:
        return np.array([(i, i) for i in range(c_mode)], dtype=np.int32).reshape(-1, 2)
    iu = np.triu_indices(c_mode)
    return np.stack(iu, 1).astype(np.int32)


def cross_correlation(eigval, eigvec, pairs, precision=2, device=None):
    """float32 [P, N, N]: out[p, a, b] = sign(s) log2(|s| + 1), s = float32(lambda_i lambda_j
    (v_i[a] . v_j[b])) * 100^precision for (i, j) = pairs[p]."""
    ops.require_gpu()
    device = device or (eigvec.device if isinstance(eigvec, torch.Tensor) and eigvec.is_cuda
                        else torch.device("cuda", torch.cuda.current_device()))
    lam, ev = _modes(eigval, eigvec, device)
    k, N = ev.shape[0], ev.shape[1]
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    lam_h = np.ascontiguousarray(lam.cpu().numpy())
    out = torch.empty((pr.shape[0], N, N), dtype=torch.float32, device=device)
    call("inch_cross_correlation", pr.shape[0], pr.ctypes.data, k, N, lam_h.ctypes.data, P(ev), int(precision),
         P(out), ops.stream_ptr())
    return out
