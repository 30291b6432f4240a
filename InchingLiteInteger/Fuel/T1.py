"""Drop-in for ``InchingLiteInteger/Fuel/T1.py`` of the reference: the ANM Hessian builder.

``Xnumpy_SparseCupyMatrixUngappped`` (reference :66-222) keeps its constructor arguments
and its methods ``ReturnNumberTotalBatch`` and ``ReturnCupyHLowerTriangleInt32/Int64``
(reference :228-660, :666-1021), which return ``(A, A_diag)``. Here ``A`` is an
``inching_amd.hessian.BlockHessian``: full-storage 3x3 block rows on the GPU that behave
like the reference's lower-triangle CSR for the operations callers use (``A @ v`` = L v,
``A.T @ v`` = L^T v, ``A.diagonal(0)``, ``A.shape``, ``A.dtype``) and that every Burn
solver of this package consumes directly. ``A_diag`` is the float64 diagonal of H + PlusI I
(a torch tensor on the GPU).

``User_DictCharmmGuiPbc`` (a CHARMM-GUI box dictionary, nm) adds the periodic x/y images of
reference :98-128, :317-396: each block sums -d d^T / |d|^2 over the images of j within the
cutoff.

As in the reference, X (nm) is expected in the order the caller wants the matrix in
(normally after ``X_KdCuthillMckeeOrder``); the builder does not reorder. The HalfNnz
buffer-size argument is checked like the reference's preallocated buffers: a matrix with
more lower-triangle scalars than ``User_MaxHalfNnzBufferSize`` raises IndexError.
"""
import numpy as np
import torch

from inching_amd import ops
from inching_amd.hessian import build_hessian


class Xnumpy_SparseCupyMatrixUngappped:
    def __init__(self, X, batch_head=None, maxleafsize=100, rc_Gamma=8.0, device=None, User_PlusI=1.0,
                 dtype_temp=torch.float64, X_precision=None, NnzMinMaxDict=None, User_DictCharmmGuiPbc=None):
        ops.require_gpu()
        self.box, self.origin = None, (0.0, 0.0)
        if User_DictCharmmGuiPbc is not None:
            from InchingLiteInteger.Fuel.Coordinate.T1 import pbc_box
            self.box, self.origin = pbc_box(User_DictCharmmGuiPbc, rc_Gamma / 10.0)
        self.X = X
        self.n_atoms = int(X.shape[0])
        self.rc_Gamma = rc_Gamma / 10.0
        self.User_PlusI = float(User_PlusI)
        self.dof = 3 * self.n_atoms
        self.device = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        if isinstance(self.device, int):
            self.device = torch.device("cuda", self.device)
        if batch_head is None:
            batch_head = torch.as_tensor(ops.batch_head(self.n_atoms, maxleafsize))
        self.batch_head = batch_head
        self.n_batch_min1 = int(len(batch_head)) - 1
        self.LeftRightNnzBound = NnzMinMaxDict
        self.dtype_temp = dtype_temp

    def ReturnNumberTotalBatch(self):
        return self.n_batch_min1 + 1

    def _build(self, User_MaxHalfNnzBufferSize, precision):
        A = build_hessian(self.X, self.rc_Gamma, plus_i=self.User_PlusI, reorder=False, precision=precision,
                          device=self.device, box=self.box, origin=self.origin)
        half = (A.nnzb - A.n_atoms) // 2 * 9 + A.n_atoms * 6
        if User_MaxHalfNnzBufferSize is not None and half > int(User_MaxHalfNnzBufferSize) + 1:
            raise IndexError("the lower triangle holds %d scalars, more than User_MaxHalfNnzBufferSize=%d"
                             % (half, int(User_MaxHalfNnzBufferSize)))
        return A, A.A_diag

    def ReturnCupyHLowerTriangleInt32(self, User_MaxHalfNnzBufferSize=1e8, precision=("f64",)):
        return self._build(User_MaxHalfNnzBufferSize, precision)

    def ReturnCupyHLowerTriangleInt64(self, User_MaxHalfNnzBufferSize=1e8, precision=("f64",)):
        return self._build(User_MaxHalfNnzBufferSize, precision)
