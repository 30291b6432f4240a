"""Drop-in for ``InchingLiteInteger/Burn/Visualisation/T1.py``: elementwise helpers of the
mode analysis, as torch expressions on the caller's tensors.

* ``S_LogModulusS(S, precision=1)`` (reference :44-52): sign(S) log2(|S 100^precision| + 1).
* ``M_BatchUnitVector`` / ``M_BatchMagnitude`` / ``M_BatchMinMaxVector`` (:58-75): per-atom
  unit vectors and magnitudes of [b, N, 3] batches.
"""
import torch


@torch.no_grad()
def S_LogModulusS(S, precision=1):
    S = S * (100.0 ** precision)
    return torch.sign(S) * torch.log2(torch.abs(S) + 1.0)


@torch.no_grad()
def M_BatchUnitVector(M, axis=2):
    return M / torch.sqrt(torch.sum(M * M, axis=axis)).unsqueeze(axis)


def M_BatchMagnitude(M, axis=2):
    return torch.sqrt(torch.sum(M * M, axis=axis)).unsqueeze(axis)


@torch.no_grad()
def M_BatchMinMaxVector(M, axis=2):
    return M_BatchUnitVector(M, axis=axis)
