"""ANM Hessian on the device: construction pipeline and the block-row matrix container.

Pipeline (replaces ``X_KdCuthillMckeeOrder`` + ``X_KdUngappedMinMaxNeighbor`` +
``Xnumpy_SparseCupyMatrixUngappped.ReturnCupyHLowerTriangleInt32`` of the reference,
``InchingLiteInteger/Fuel/Coordinate/T1.py:171-481`` and ``InchingLiteInteger/Fuel/T1.py:67-660``):

1. cell-list neighbour CSR of the input coordinates (device kernel);
2. reverse Cuthill-McKee order (host C++, bit-exact to the reference's 3DRCM with a
   pinned stable tie rule, or any supplied seed order);
3. coordinates permuted on the device, neighbour CSR recomputed in the new order;
4. 3x3 super-element assembly straight into full-storage block rows, in float64 and/or
   float32 (the float32 copy is the round-to-nearest of the float64 values).

The container ``BlockHessian`` keeps the coordinates and the neighbour pattern, so the
matrix-free FP64 operator (residual polish) needs no float64 copy of the values.
"""
import numpy as np
import torch

from . import ops


class BlockHessian:
    """A = H + PlusI * I, full-storage 3x3 block rows in RCM order on the GPU.

    Quacks like the reference's lower-triangle CSR where callers rely on it:
    ``A @ v`` is the scalar lower triangle L v, ``A.T @ v`` is L^T v and
    ``A.diagonal(0)`` is A_diag, so ``A @ v + A.T @ v - A_diag * v == full A v`` as in
    ``Burn/Krylov/T3.py:131-160``. ``A.matvec(v)`` is the fused full product.
    """

    def __init__(self, xyz, rowptr, col, vals64, vals32, diag, plus_i, rc, order=None, undo=None, images=None,
                 box=None):
        self.images = images        # periodic-image masks per entry (None: open boundaries)
        self.box = box              # (box_x, box_y) of the periodic images
        self.xyz = xyz
        self.rowptr = rowptr
        self.col = col
        self.vals64 = vals64
        self.vals32 = vals32
        self.A_diag = diag
        self.plus_i = float(plus_i)
        self.rc = float(rc)
        self.order = order
        self.undo = undo
        self.n_atoms = int(xyz.shape[0])
        self.shape = (3 * self.n_atoms, 3 * self.n_atoms)
        self.ndim = 2
        self.lanes = 16
        self._transposed = False

    # ----------------------------------------------------------- properties --
    @property
    def dtype(self):
        return np.dtype(np.float64) if self.vals64 is not None else np.dtype(np.float32)

    @property
    def device(self):
        return self.xyz.device

    @property
    def nnzb(self):
        return int(self.col.numel())

    def vals(self, precision):
        v = self.vals64 if precision == "f64" else self.vals32
        if v is None:
            raise ValueError("this matrix holds no %s values" % precision)
        return v

    def col16(self):
        """(rowbase, col16) 16-bit column offsets of the pattern (built once), or None when a
        row spans too many columns for 16 bits."""
        if "_col16" not in self.__dict__:
            try:
                self._col16 = ops.bsr_col16(self.n_atoms, self.rowptr, self.col)
            except ops.InchingError:
                self._col16 = None
        return self._col16

    def clusters(self, cell_xy=None, cell_z=None):
        """Cluster layout of the matrix-free operator (ops.anm_clusters; built once), or None for
        periodic matrices and when the layout does not fit. Cells default to 0.5 rc across and
        0.25 rc along z, which for protein densities gives clusters of 4 atoms a union of about
        1.6 times one atom's neighbours."""
        if "_clusters" not in self.__dict__:
            self._clusters = None
            if self.images is None:
                try:
                    self._clusters = ops.anm_clusters(self.xyz, self.rowptr, self.col,
                                                      cell_xy or 0.5 * self.rc, cell_z or 0.25 * self.rc)
                except ops.InchingError:
                    self._clusters = None
        return self._clusters

    def records(self):
        """Record form of the cluster layout for the Chebyshev recurrence (ops.anm_records:
        three record buffers and the lists in slot numbers; built once), or None when the
        cluster layout is unavailable or too large for 32-bit record offsets."""
        if "_records" not in self.__dict__:
            self._records = None
            cl = self.clusters()
            if cl is not None:
                try:
                    self._records = ops.anm_records(self.xyz, cl)
                except ops.InchingError:
                    self._records = None
        return self._records

    def ensure(self, precision):
        """Make sure the values of ``precision`` exist (assembled on demand)."""
        if (precision == "f64" and self.vals64 is None) or (precision == "f32" and self.vals32 is None):
            if self.images is not None:
                v64, v32, _ = ops.hessian_assemble_pbc(self.xyz, self.plus_i, self.rowptr, self.col, self.images,
                                                       self.box, want64=precision == "f64",
                                                       want32=precision == "f32")
            else:
                v64, v32, _ = ops.hessian_assemble(self.xyz, self.rc, self.plus_i, self.rowptr, self.col,
                                                   want64=precision == "f64", want32=precision == "f32")
            if precision == "f64":
                self.vals64 = v64
            else:
                self.vals32 = v32

    # ------------------------------------------------------------ products --
    def matvec(self, x, y=None, precision=None, **epi):
        """y = A x (+ fused epilogue, see ops.bsr_spmv); precision 'f64' / 'f32'."""
        if precision is None:
            precision = "f64" if self.vals64 is not None else "f32"
        if y is None:
            y = torch.empty_like(x)
        ops.bsr_spmv(self.n_atoms, self.rowptr, self.col, self.vals(precision), x, y, lanes=self.lanes, **epi)
        return y

    def matmat(self, X, precision=None):
        if precision is None:
            precision = "f64" if self.vals64 is not None else "f32"
        Y = torch.empty_like(X)
        ops.bsr_spmm(self.n_atoms, self.rowptr, self.col, self.vals(precision), X.contiguous(), Y)
        return Y

    def matfree_spmv(self, x, y, config=0, **epi):
        """Matrix-free float64 product with the fused epilogue (``config``: autotuner code
        lanes*10+unroll, plus 1000 for 16-bit column offsets, 2000 for XCD-contiguous rows and
        4000 for the cluster layout, 0 = default; periodic matrices use their own kernel)."""
        if self.images is not None:
            ops.matfree_pbc(self.xyz, self.rowptr, self.col, self.images, self.box, self.plus_i, x, y, **epi)
            return
        shape, flags = config % 1000, config // 1000
        abi = shape + (2000 if flags & 2 else 0) if shape else 0
        if flags & 4:
            cl = self.clusters()
            if cl is None:
                raise ValueError("config %d asks for the cluster layout, which this matrix cannot use" % config)
            ops.matfree_cl(self.xyz, cl, self.plus_i, x, y, config=abi, **epi)
            return
        c16 = self.col16() if flags & 1 else None
        if flags & 1 and c16 is None:
            raise ValueError("config %d asks for 16-bit column offsets, which this matrix cannot use" % config)
        ops.matfree_spmv(self.xyz, self.rowptr, self.col, self.plus_i, x, y, config=abi, col16=c16, **epi)

    def matfree(self, X):
        """FP64 (H + PlusI I) X recomputed from coordinates (X [3n] or [3n, r])."""
        Y = torch.empty_like(X)
        if self.images is not None:
            ops.matfree_pbc(self.xyz, self.rowptr, self.col, self.images, self.box, self.plus_i, X.contiguous(), Y)
        elif self.clusters() is not None:
            ops.matfree_cl_apply(self.xyz, self.clusters(), self.plus_i, X.contiguous(), Y)
        else:
            ops.matfree_apply(self.xyz, self.rowptr, self.col, self.plus_i, X.contiguous(), Y)
        return Y

    def _part(self, v, part):
        prec = "f64" if self.vals64 is not None else "f32"
        vv = v.reshape(-1).contiguous()
        if prec == "f64" and vv.dtype != torch.float64:
            vv = vv.double()
        y = torch.empty_like(vv)
        ops.bsr_spmv_part(self.n_atoms, self.rowptr, self.col, self.vals(prec), vv, y, part)
        return y.reshape(v.shape)

    def __matmul__(self, v):
        if v.dim() == 2 and v.shape[1] != 1:
            return torch.stack([self.__matmul__(v[:, r]) for r in range(v.shape[1])], dim=1)
        return self._part(v, 2 if self._transposed else 1)

    @property
    def T(self):
        t = BlockHessian.__new__(BlockHessian)
        t.__dict__.update(self.__dict__)
        t._transposed = not self._transposed
        return t

    def diagonal(self, k=0):
        if k != 0:
            raise NotImplementedError("only the main diagonal is available")
        return self.A_diag

    # ------------------------------------------------------- reference CSR --
    def to_lower_csr(self):
        """Scalar CSR (indptr, indices, data) of tril(A) in float64 numpy arrays (host)."""
        import scipy.sparse as sp
        self.ensure("f64")
        rp = self.rowptr.cpu().numpy()
        col = self.col.cpu().numpy().astype(np.int64)
        vals = self.vals64.cpu().numpy().reshape(-1, 3, 3)
        B = sp.bsr_matrix((vals, col, rp), shape=self.shape).tocsr()
        L = sp.tril(B, k=0, format="csr")
        L.eliminate_zeros()
        L.sort_indices()
        return L.indptr.astype(np.int64), L.indices.astype(np.int64), L.data


def _to_device_xyz(X, device):
    if isinstance(X, torch.Tensor):
        t = X.detach().to(device=device, dtype=torch.float64)
    else:
        t = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float64), device=device)
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError("coordinates must be [n_atoms, 3]")
    return t.contiguous()


def rcm_order(X, rc, seed_order=None, reverse=True, device=None, rcm="device"):
    """3DRCM permutation of coordinates X: (order int32, undo int64) numpy arrays.
    ``rcm``: 'device' (level-synchronous BFS kernels) or 'host' (sequential C++); both give
    the same permutation."""
    ops.require_gpu()
    device = device or torch.device("cuda", torch.cuda.current_device())
    xyz = _to_device_xyz(X, device)
    indptr, indices = ops.neighbours(xyz, rc)
    if rcm == "device":
        order, undo = ops.rcm_device(indptr, indices, seed_order=seed_order, reverse=reverse)
        return order.cpu().numpy(), undo.cpu().numpy()
    return ops.rcm_host(indptr.cpu().numpy(), indices.cpu().numpy(), seed_order=seed_order, reverse=reverse)


def build_hessian(X, rc, plus_i=1.0, reorder=True, precision=("f32",), seed_order=None, device=None, rcm="device",
                  box=None, origin=(0.0, 0.0)):
    """Coordinates -> BlockHessian (in RCM order when ``reorder``).

    ``precision``: which value copies to assemble ('f32', 'f64'). Returns the container;
    ``A.order`` / ``A.undo`` hold the permutation (identity when not reordered). ``rcm``:
    'device' or 'host' ordering (the same permutation). ``box`` = (box_x, box_y): periodic
    images in x and y (box corner ``origin``); the ordering then runs on the periodic
    neighbour graph.
    """
    ops.require_gpu()
    device = device or torch.device("cuda", torch.cuda.current_device())
    xyz = _to_device_xyz(X, device)
    n = xyz.shape[0]

    def pattern(xyz_):
        if box is None:
            return ops.neighbours(xyz_, rc) + (None,)
        return ops.neighbours_pbc(xyz_, rc, box, origin)

    comps = None                     # connected components (offsets in the RCM order)
    permuted = None                  # the search's CSR renumbered into the RCM order
    if reorder:
        indptr, indices, _ = pattern(xyz)
        if rcm == "device":
            order_d, undo_d, comps = ops.rcm_device(indptr, indices, seed_order=seed_order, components=True)
            xyz = xyz[order_d.long()].contiguous()
            order, undo = order_d.cpu().numpy(), undo_d.cpu().numpy()
            if box is None:
                permuted = ops.csr_permute(indptr, indices, order_d, undo_d)
        else:
            order, undo = ops.rcm_host(indptr.cpu().numpy(), indices.cpu().numpy(), seed_order=seed_order)
            xyz = xyz[torch.as_tensor(order.astype(np.int64), device=device)].contiguous()
        del indptr, indices
    else:
        order = np.arange(n, dtype=np.int32)
        undo = np.arange(n, dtype=np.int64)
    if permuted is not None:
        (rowptr, col), images = permuted, None
    else:
        rowptr, col, images = pattern(xyz)
    if images is None:
        v64, v32, diag = ops.hessian_assemble(xyz, rc, plus_i, rowptr, col, want64="f64" in precision,
                                              want32="f32" in precision)
    else:
        v64, v32, diag = ops.hessian_assemble_pbc(xyz, plus_i, rowptr, col, images, box, want64="f64" in precision,
                                                  want32="f32" in precision)
    A = BlockHessian(xyz, rowptr, col, v64, v32, diag, plus_i, rc, order=order, undo=undo, images=images,
                     box=None if box is None else (float(box[0]), float(box[1])))
    A.components = comps
    return A


def rigid_body_basis(xyz, rotations=True):
    """6 x 3N rigid-body basis of device coordinates, built as the reference driver builds it
    (Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:314-372): three unit
    translations (1/sqrt N), three infinitesimal rotations X @ R_a normalised, then two sweeps
    of double Gram-Schmidt. float64 torch [6, 3N] on the device of ``xyz``. With periodic
    images only the translations are zero modes (``rotations=False``: the 3 x 3N block)."""
    x = xyz.to(torch.float64)
    n = x.shape[0]
    Q = torch.zeros((6, 3 * n), dtype=torch.float64, device=x.device)
    for a in range(3):
        Q[a].view(n, 3)[:, a] = 1.0 / np.sqrt(n)
    # e_a x p for the three axes
    rots = [torch.stack([torch.zeros_like(x[:, 0]), -x[:, 2], x[:, 1]], 1),
            torch.stack([x[:, 2], torch.zeros_like(x[:, 0]), -x[:, 0]], 1),
            torch.stack([-x[:, 1], x[:, 0], torch.zeros_like(x[:, 0])], 1)]
    for a in range(3):
        g = rots[a].reshape(-1)
        Q[3 + a] = g / torch.linalg.norm(g)
    for _ in range(2):
        for ix in range(1, 6):
            for _rep in range(2):
                Q[ix] -= Q[:ix].T @ (Q[:ix] @ Q[ix])
                Q[ix] /= torch.sqrt(Q[ix] @ Q[ix])
    return Q if rotations else Q[:3].contiguous()


def component_rigid_basis(xyz, comp):
    """Rigid-body basis per connected component (6 rows per component)."""
    comp = torch.as_tensor(comp, device=xyz.device)
    rows = []
    for c in torch.unique(comp).tolist():
        idx = torch.nonzero(comp == c).flatten()
        Qc = rigid_body_basis(xyz[idx])
        full = torch.zeros((6, 3 * xyz.shape[0]), dtype=torch.float64, device=xyz.device)
        dof = (3 * idx[:, None] + torch.arange(3, device=xyz.device)[None, :]).reshape(-1)
        full[:, dof] = Qc
        rows.append(full)
    return torch.cat(rows, 0)
