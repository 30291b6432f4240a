"""Drop-in for ``InchingLiteInteger/Burn/Visualisation/T2.py``: square fluctuations and
cross-correlation maps of computed modes, each one kernel of the library
(``inching_amd.analysis``).

* ``Heigval_Heigvec_SqFluctPerMode_SqFluct_RatioSqFPM_RatioVar(H_eigval, H_eigvec)``
  (reference :43-55) returns (SqFluctPerMode [k, N, 1], SqFluct [N], RatioSqFPM [k, N, 1],
  RatioVariance [k]).
* ``Heigval_Heigvec_HccPairlist_HccBatch(H_eigval, H_eigvec, plot_CC=True, device=..., c_mode=5,
  SelfOnly=True)`` (reference :58-160) returns (``torch.triu_indices(c_mode, c_mode).T``, the
  float32 [P, N, N] log-modulus maps of the pairs (i, i), i < c_mode, or of every (i, j),
  i <= j, with ``SelfOnly=False``). The reference's first return value is the full upper
  triangle whatever ``SelfOnly`` says; that is kept. ``plot_CC`` drew an image grid with
  matplotlib; here no figure is drawn (the maps are returned either way).

H_eigvec is [k, N, 3] (mode, atom, xyz); numpy or torch, results are device tensors (float64
for the fluctuations, or the input's float32 when the modes are float32).
"""
import torch

from inching_amd import analysis


@torch.no_grad()
def Heigval_Heigvec_SqFluctPerMode_SqFluct_RatioSqFPM_RatioVar(H_eigval, H_eigvec):
    sfpm, sqf, ratio, rv = analysis.square_fluctuations(H_eigval, H_eigvec)
    if isinstance(H_eigvec, torch.Tensor) and H_eigvec.dtype == torch.float32:
        sfpm, sqf, ratio, rv = sfpm.float(), sqf.float(), ratio.float(), rv.float()
    return sfpm.unsqueeze(2), sqf, ratio.unsqueeze(2), rv


@torch.no_grad()
def Heigval_Heigvec_HccPairlist_HccBatch(H_eigval, H_eigvec, plot_CC=True, device=None, c_mode=5, SelfOnly=True):
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    pairs = analysis.mode_pairs(c_mode, self_only=SelfOnly)
    hcc = analysis.cross_correlation(H_eigval, H_eigvec, pairs, precision=2, device=torch.device(device))
    return torch.triu_indices(c_mode, c_mode).T, hcc
