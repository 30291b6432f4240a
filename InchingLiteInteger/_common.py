"""Helpers shared by the shim modules: tensors in/out, dense Q handling, device scalars."""
import numpy as np
import torch


class DevScalar(float):
    """A float that also answers ``.get()`` (the reference returns 0-d device arrays whose
    callers call ``.get()``, e.g. the spectrum bound in the CTRLM notebooks)."""

    def get(self):
        return float(self)


def device_of(A):
    return getattr(A, "device", torch.device("cuda", torch.cuda.current_device()))


def as_vec(v, device, dtype=torch.float64):
    if isinstance(v, torch.Tensor):
        t = v.to(device=device, dtype=dtype)
    else:
        t = torch.as_tensor(np.asarray(v), device=device, dtype=dtype)
    return t.reshape(-1).contiguous()


def dense_Q(Q, device):
    """The rigid-body basis as float64 [m, n] rows on the device (dense or sparse input)."""
    if Q is None:
        return None
    if isinstance(Q, torch.Tensor):
        if Q.is_sparse or Q.layout != torch.strided:
            Q = Q.to_dense()
        return Q.to(device=device, dtype=torch.float64).contiguous()
    for attr in ("toarray", "todense"):
        if hasattr(Q, attr):
            Q = getattr(Q, attr)()
            break
    if hasattr(Q, "get"):
        Q = Q.get()
    return torch.as_tensor(np.asarray(Q, dtype=np.float64), device=device).contiguous()


def require_block_matrix(A):
    from inching_amd.hessian import BlockHessian
    if not isinstance(A, BlockHessian):
        raise TypeError("this implementation operates on the matrix returned by "
                        "Xnumpy_SparseCupyMatrixUngappped.ReturnCupyHLowerTriangleInt32/Int64 "
                        "(got %s)" % type(A).__name__)
    return A
