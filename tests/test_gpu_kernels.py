"""GPU parity of the construction and operator kernels against the CPU oracle.

Integer/index outputs (neighbour CSR, RCM order) must be bit-exact; float64 values agree
to rounding (the diagonal sums run in a different order); float32 values are the
round-to-nearest of the float64 values (<= 1 ulp); float32 SpMV has relative error
<= 1e-6 in the 2-norm.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import geometry as OG
from oracle import hessian as OH
from oracle import rcm as OR
from tests.helpers import projected_system, small_globule

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


@pytest.fixture(scope="module")
def glob2k(cuda):
    X = small_globule(2000, seed=21)
    return X


@pytest.fixture(scope="module")
def glob40k(cuda):
    from inching_amd import synth
    return synth.globule(40000, 3)


def _print(*a):
    print(*a, flush=True)


def test_neighbours_bit_exact_small(glob2k):
    from inching_amd import ops
    _print("case: neighbours n=2000 rc=8")
    ip_ref, ix_ref = OG.neighbour_csr(glob2k, 8.0)
    ip, ix = ops.neighbours(_dev(glob2k), 8.0)
    assert np.array_equal(ip.cpu().numpy(), ip_ref)
    assert np.array_equal(ix.cpu().numpy(), ix_ref)


def test_neighbours_bit_exact_40k_nm_units(glob40k):
    from inching_amd import ops
    _print("case: neighbours n=40000 in nm, rc=0.8")
    Xnm = glob40k / 10.0
    ip_ref, ix_ref = OG.neighbour_csr(Xnm, 0.8)
    ip, ix = ops.neighbours(_dev(Xnm), 0.8)
    assert np.array_equal(ip.cpu().numpy(), ip_ref)
    assert np.array_equal(ix.cpu().numpy(), ix_ref)


def test_neighbours_long_rows_and_renumbering(glob2k):
    """Rows of several hundred entries (rc = 20 on 2,000 atoms: up to ~1,000 neighbours), which
    the row sort handles by its bitonic path (the rank path stops at 256), bit-exact against
    the oracle; and the CSR renumbered into an RCM order equals the oracle's CSR of the
    reordered coordinates."""
    from inching_amd import ops
    _print("case: neighbours n=2000 rc=20 (long rows)")
    ip_ref, ix_ref = OG.neighbour_csr(glob2k, 20.0)
    assert np.diff(ip_ref).max() > 256
    ip, ix = ops.neighbours(_dev(glob2k), 20.0)
    assert np.array_equal(ip.cpu().numpy(), ip_ref)
    assert np.array_equal(ix.cpu().numpy(), ix_ref)
    order, undo = ops.rcm_device(ip, ix)
    rp, rc_ = ops.csr_permute(ip, ix, order, undo)
    o = order.cpu().numpy()
    ip2, ix2 = OG.neighbour_csr(glob2k[o], 20.0)
    assert np.array_equal(rp.cpu().numpy(), ip2) and np.array_equal(rc_.cpu().numpy(), ix2)


def test_cluster_layout_rejects_unsorted_rows(glob2k):
    """The cluster lists merge the atoms' rows, which must be in ascending column order with
    columns in [0, n): a row out of order or a column out of range is an error
    (INCH_ERR_INVALID), found before any merge rank or coordinate is formed from it."""
    from inching_amd import ops
    from inching_amd._lib import InchingError
    _print("case: cluster layout, unsorted rows")
    ip, ix = ops.neighbours(_dev(glob2k), 8.0)
    bad = ix.clone()
    b0 = int(ip[5].item())
    bad[b0], bad[b0 + 1] = ix[b0 + 1], ix[b0]          # swap two columns of row 5
    xyz = _dev(glob2k)
    with pytest.raises(InchingError):
        ops.anm_clusters(xyz, ip, bad, 4.0, 2.0)
    far = ix.clone()
    far[int(ip[6].item()) - 1] = len(glob2k) + 5        # last column of row 5 beyond n (still ascending)
    with pytest.raises(InchingError):
        ops.anm_clusters(xyz, ip, far, 4.0, 2.0)
    ops.anm_clusters(xyz, ip, ix, 4.0, 2.0)             # the sorted rows are accepted


def test_rcm_pipeline_matches_oracle(glob2k):
    from inching_amd.hessian import rcm_order
    _print("case: rcm pipeline n=2000")
    ip, ix = OG.neighbour_csr(glob2k, 8.0)
    o_ref, u_ref = OR.rcm_order(ip, ix)
    o, u = rcm_order(glob2k, 8.0)
    assert np.array_equal(o, o_ref) and np.array_equal(u, u_ref)


@pytest.mark.parametrize("case", ["globule", "blobs_and_singletons", "blobs_random_seeds", "many_blobs",
                                  "numpy_default_seeds", "forward", "40k"])
def test_rcm_device_bit_exact(glob2k, glob40k, case):
    """The level-synchronous device RCM gives the oracle's permutation (and, at 40k atoms,
    the sequential host implementation's), for several components and isolated atoms, an
    explicit seed order (also a random one, which interleaves the components' first seeds),
    40 components of different depths advancing together, and the unreversed order."""
    from inching_amd import ops
    _print("case: device rcm %s" % case)
    rng = np.random.default_rng(5)
    if case in ("blobs_and_singletons", "blobs_random_seeds"):
        X = np.concatenate([rng.uniform(0, 15, (300, 3)), rng.uniform(60, 70, (150, 3)),
                            np.array([[200.0, 0, 0], [-200.0, 0, 0], [0, 300.0, 0]])])
    elif case == "many_blobs":
        X = np.concatenate([rng.uniform(0, 6 + 2 * (b % 7), (40 + 13 * b, 3)) + np.array([100.0 * b, 0, 0])
                            for b in range(40)])
    elif case == "40k":
        X = glob40k
    else:
        X = glob2k
    xyz = _dev(X)
    ip_d, ix_d = ops.neighbours(xyz, 8.0)
    ip, ix = ip_d.cpu().numpy(), ix_d.cpu().numpy()
    seeds = np.argsort(np.diff(ip)) if case == "numpy_default_seeds" else None
    if case == "blobs_random_seeds":
        seeds = rng.permutation(len(ip) - 1)
    reverse = case != "forward"
    o, u, comps = ops.rcm_device(ip_d, ix_d, seed_order=seeds, reverse=reverse, components=True)
    o, u = o.cpu().numpy(), u.cpu().numpy()
    # components are contiguous in the order and their offsets match scipy's labelling
    from scipy.sparse.csgraph import connected_components
    nc, lab = connected_components(sp.csr_matrix((np.ones(ix.size), ix, ip), shape=(len(ip) - 1,) * 2),
                                   directed=False)
    lab_o = lab[o]
    want = np.r_[np.nonzero(np.r_[True, lab_o[1:] != lab_o[:-1]])[0], len(o)]
    assert len(want) == nc + 1 and np.array_equal(comps, want)
    if case in ("blobs_and_singletons", "blobs_random_seeds"):
        assert nc == 5
    if case == "many_blobs":
        assert nc == 40
    if case == "40k":
        o_ref, u_ref = ops.rcm_host(ip, ix)
    else:
        o_ref, u_ref = OR.rcm_order(ip, ix, seed_order=seeds, reverse=reverse)
    assert np.array_equal(o, o_ref) and np.array_equal(u, u_ref)


def test_rcm_device_rejects_bad_input(glob2k):
    from inching_amd import ops
    from inching_amd._lib import InchingError
    _print("case: device rcm bad input")
    ip_d, ix_d = ops.neighbours(_dev(glob2k), 8.0)
    n = ip_d.numel() - 1
    bad = np.arange(n)
    bad[5] = n + 7                                   # out of range: caught before any indexed access
    with pytest.raises(InchingError):
        ops.rcm_device(ip_d, ix_d, seed_order=bad)
    dup = np.arange(n)
    dup[5] = dup[6]                                  # not a permutation
    with pytest.raises(InchingError):
        ops.rcm_device(ip_d, ix_d, seed_order=dup)


def test_assembly_matches_oracle(glob2k):
    from inching_amd.hessian import build_hessian
    _print("case: assembly n=2000")
    A = build_hessian(glob2k, 8.0, plus_i=1.0, reorder=True, precision=("f32", "f64"))
    Xp = glob2k[A.order]
    rp_ref, col_ref, v_ref = OH.block_hessian(Xp, 8.0, 1.0)
    assert np.array_equal(A.rowptr.cpu().numpy(), rp_ref)
    assert np.array_equal(A.col.cpu().numpy(), col_ref)
    v64 = A.vals64.cpu().numpy().reshape(-1, 3, 3)
    assert np.max(np.abs(v64 - v_ref)) <= 1e-13 * max(1.0, np.max(np.abs(v_ref)))
    v32 = A.vals32.cpu().numpy().reshape(-1, 3, 3)
    ulp = np.spacing(np.abs(v_ref).astype(np.float32))
    assert np.all(np.abs(v32.astype(np.float64) - v_ref) <= ulp)
    d_ref = np.concatenate([np.diag(v) for v in v_ref[col_ref == np.repeat(np.arange(2000), np.diff(rp_ref))]])
    assert np.allclose(A.A_diag.cpu().numpy(), d_ref, rtol=1e-13, atol=1e-13)


def test_assembly_invariants_40k(glob40k):
    from inching_amd.hessian import build_hessian, rigid_body_basis
    _print("case: assembly invariants n=40000")
    A = build_hessian(glob40k, 8.0, plus_i=1.0, reorder=True, precision=("f64",))
    Q = rigid_body_basis(A.xyz)
    AQ = A.matmat(Q.T.contiguous(), precision="f64")
    # rigid-body modes are eigenvectors of A = H + I with eigenvalue 1
    assert torch.max(torch.abs(AQ - Q.T)).item() < 1e-10
    # trace identity: tr(H) = 2 * #pairs
    n_pairs = (A.nnzb - A.n_atoms) // 2
    assert abs((A.A_diag.sum().item() - 3 * A.n_atoms) - 2 * n_pairs) < 1e-6 * n_pairs


@pytest.fixture(scope="module")
def mat(glob2k):
    from inching_amd.hessian import build_hessian
    A = build_hessian(glob2k, 8.0, plus_i=1.0, reorder=True, precision=("f32", "f64"))
    rp = A.rowptr.cpu().numpy()
    col = A.col.cpu().numpy()
    Bs = sp.bsr_matrix((A.vals64.cpu().numpy().reshape(-1, 3, 3), col.astype(np.int64), rp), shape=A.shape).tocsr()
    return A, Bs


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_spmv_f64_f32(mat, lanes):
    from inching_amd import ops
    A, Bs = mat
    _print("case: spmv lanes=%d" % lanes)
    x = np.random.default_rng(lanes).standard_normal(A.shape[0])
    ref = Bs @ x
    y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals64, _dev(x), y, lanes=lanes)
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-13 * np.linalg.norm(ref)
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, _dev(x), y, lanes=lanes)   # f32 matrix, f64 vectors
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-6 * np.linalg.norm(ref)
    y32 = torch.empty(A.shape[0], dtype=torch.float32, device="cuda")
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, _dev(x, torch.float32), y32, lanes=lanes)
    assert np.linalg.norm(y32.cpu().numpy().astype(np.float64) - ref) <= 1e-6 * np.linalg.norm(ref)


def test_spmv_deterministic(mat):
    from inching_amd import ops
    A, _ = mat
    x = _dev(np.random.default_rng(0).standard_normal(A.shape[0]), torch.float32)
    y1 = torch.empty_like(x)
    y2 = torch.empty_like(x)
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, x, y1)
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, x, y2)
    assert torch.equal(y1, y2)


def test_spmv_epilogues(mat):
    from inching_amd import ops
    from inching_amd.hessian import rigid_body_basis
    A, Bs = mat
    _print("case: spmv epilogues")
    rng = np.random.default_rng(3)
    n = A.shape[0]
    x, z, acc = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    s1, c0, c2, mu = 0.07, 12.5, 1.0, -0.3
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    accd = _dev(acc)
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals64, _dev(x), y, s1=s1, c0=c0, z=_dev(z), c2=c2, acc=accd, mu=mu)
    out = s1 * (Bs @ x - c0 * x) - c2 * z
    assert np.allclose(y.cpu().numpy(), out, rtol=1e-12, atol=1e-12)
    assert np.allclose(accd.cpu().numpy(), acc + mu * out, rtol=1e-12, atol=1e-12)
    Q = rigid_body_basis(A.xyz)
    qc = _dev(rng.standard_normal(6))
    ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals64, _dev(x), y, Q=Q, qcoef=qc)
    ref = Bs @ x + Q.cpu().numpy().T @ qc.cpu().numpy()
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


def test_reference_half_storage_views(mat):
    A, Bs = mat
    x = np.random.default_rng(4).standard_normal(A.shape[0])
    L = sp.tril(Bs, format="csr")
    xd = _dev(x)
    assert np.allclose((A @ xd).cpu().numpy(), L @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose((A.T @ xd).cpu().numpy(), L.T @ x, rtol=1e-13, atol=1e-13)
    full = (A @ xd) + (A.T @ xd) - A.diagonal(0) * xd
    assert np.allclose(full.cpu().numpy(), Bs @ x, rtol=1e-12, atol=1e-12)


def test_spmm_and_matfree(mat):
    A, Bs = mat
    X = np.random.default_rng(5).standard_normal((A.shape[0], 11))
    Y = A.matmat(_dev(X), precision="f64").cpu().numpy()
    assert np.allclose(Y, Bs @ X, rtol=1e-12, atol=1e-12)
    Y32 = A.matmat(_dev(X, torch.float32), precision="f32").cpu().numpy()
    assert np.linalg.norm(Y32 - Bs @ X) <= 1e-6 * np.linalg.norm(Bs @ X)
    Ym = A.matfree(_dev(X)).cpu().numpy()
    assert np.linalg.norm(Ym - Bs @ X) <= 1e-13 * np.linalg.norm(Bs @ X)


def test_dense_kernels(cuda):
    from inching_amd import ops
    rng = np.random.default_rng(6)
    n, J = 100003, 37
    V = rng.standard_normal((J + 3, n))
    u = rng.standard_normal(n)
    Vd, ud = _dev(V), _dev(u)
    ws = ops.Workspace(64, Vd.device)
    h = torch.empty(64, dtype=torch.float64, device="cuda")
    ops.multidot(Vd, J, ud, h, ws)
    assert np.allclose(h[:J].cpu().numpy(), V[:J] @ u, rtol=1e-12, atol=1e-9)
    nrm = torch.empty(1, dtype=torch.float64, device="cuda")
    ops.gemv_update(Vd, J, h, ud, 0.5, nrm2_out=nrm, ws=ws)
    ref = u - 0.5 * V[:J].T @ (V[:J] @ u)
    assert np.allclose(ud.cpu().numpy(), ref, rtol=1e-11, atol=1e-9)
    assert np.isclose(nrm.item(), np.linalg.norm(ref), rtol=1e-12)
    W = rng.standard_normal((J, 21))
    out = torch.empty((21, n), dtype=torch.float64, device="cuda")
    ops.rr_combine(Vd, J, _dev(W), out)
    assert np.allclose(out.cpu().numpy(), W.T @ V[:J], rtol=1e-11, atol=1e-9)
    W2 = rng.standard_normal((J, 5))
    out2 = torch.empty((5, n), dtype=torch.float64, device="cuda")
    ops.rr_combine(Vd, J, _dev(W2), out2)
    assert np.allclose(out2.cpu().numpy(), W2.T @ V[:J], rtol=1e-11, atol=1e-9)
    # more than 64 rows (two LDS chunks of W) and more than 64 outputs (two output blocks)
    Vb = rng.standard_normal((150, 20011))
    W3 = rng.standard_normal((150, 90))
    out3 = torch.empty((90, 20011), dtype=torch.float64, device="cuda")
    ops.rr_combine(_dev(Vb), 150, _dev(W3), out3)
    assert np.allclose(out3.cpu().numpy(), W3.T @ Vb, rtol=1e-11, atol=1e-9)
    dst, dst2 = torch.empty(n, dtype=torch.float64, device="cuda"), torch.empty(n, dtype=torch.float64, device="cuda")
    ops.scale_div(ud, nrm, dst, dst2)
    assert torch.allclose(dst, ud / nrm) and torch.equal(dst, dst2)
    a = _dev(np.array([0.25]))
    y = _dev(u)
    ops.axpy_dev(a, -1.0, Vd[0], y)
    assert np.allclose(y.cpu().numpy(), u - 0.25 * V[0])
    y32 = torch.empty(n, dtype=torch.float32, device="cuda")
    ops.axpby(2.0, Vd[1], 0.0, y32)
    assert np.allclose(y32.cpu().numpy(), (2.0 * V[1]).astype(np.float32))
    ops.convert(y32, dst)
    assert np.array_equal(dst.cpu().numpy(), y32.cpu().numpy().astype(np.float64))


@pytest.mark.parametrize("k1,k2,n", [(64, 64, 100003), (7, 70, 5000), (130, 6, 33333), (1, 1, 17)])
def test_gram_mfma_and_residual_norms(cuda, k1, k2, n):
    from inching_amd import ops
    print("case: gram k1=%d k2=%d n=%d" % (k1, k2, n), flush=True)
    rng = np.random.default_rng(k1 * 1000 + k2)
    X = rng.standard_normal((k1, n))
    Y = rng.standard_normal((k2, n))
    G = ops.gram(_dev(X), _dev(Y)).cpu().numpy()
    ref = X @ Y.T
    assert np.allclose(G, ref, rtol=1e-11, atol=1e-9 * np.sqrt(n))
    # strided row views (leading dimension > row length)
    Xb = _dev(np.concatenate([X, X], axis=1))
    G2 = ops.gram(Xb[:, :n], Xb[:, :n]).cpu().numpy()
    assert np.allclose(G2, X @ X.T, rtol=1e-11, atol=1e-9 * np.sqrt(n))
    th = rng.standard_normal(k1)
    Yk = rng.standard_normal((k1, n))
    out = ops.residual_norms(_dev(X), _dev(Yk), _dev(th)).cpu().numpy()
    assert np.allclose(out, np.linalg.norm(Yk - th[:, None] * X, axis=1), rtol=1e-12)


def test_matfree_multi_rhs_chunks(mat):
    A, Bs = mat
    for r in (1, 8, 13, 64):
        X = np.random.default_rng(r).standard_normal((A.shape[0], r))
        Y = A.matfree(_dev(X)).cpu().numpy()
        assert np.linalg.norm(Y - Bs @ X) <= 1e-13 * np.linalg.norm(Bs @ X)


@pytest.mark.parametrize("J,n", [(1, 1000), (37, 100003), (128, 90001), (256, 20000)])
def test_cgs2_reorth_fused(cuda, J, n):
    from inching_amd import ops
    print("case: cgs2 J=%d n=%d" % (J, n), flush=True)
    rng = np.random.default_rng(J)
    V = np.linalg.qr(rng.standard_normal((n, J)))[0].T.copy()      # orthonormal rows
    u = rng.standard_normal(n) + 3.0 * V[0]
    Vd, ud = _dev(np.concatenate([V, np.zeros((1, n))])), _dev(u)
    h = torch.empty(J, dtype=torch.float64, device="cuda")
    ws = ops.Workspace(J, Vd.device)
    ops.multidot(Vd, J, ud, h, ws)
    beta = torch.empty(1, dtype=torch.float64, device="cuda")
    ops.cgs2_reorth(Vd, J, h, ud, beta, vnext=Vd[J])
    u1 = u - V.T @ (V @ u)
    b_ref = np.linalg.norm(u1)
    u2 = u1 - V.T @ (V @ u1)
    assert np.isclose(beta.item(), b_ref, rtol=1e-12)
    assert np.allclose(ud.cpu().numpy(), u2, rtol=1e-10, atol=1e-12)
    assert np.allclose(Vd[J].cpu().numpy(), u2 / b_ref, rtol=1e-10, atol=1e-12)
    assert np.max(np.abs(V @ Vd[J].cpu().numpy())) < 1e-13


@pytest.mark.parametrize("i,kk,eps", [(0, 0, 1e-4), (1, 0, 1e-4), (20, 0, 1e-4), (20, 6, 1e-4), (20, 6, 0.0),
                                      (40, 39, 1e-4), (300, 100, 1e-5)])
def test_dcgs2_step_first_order_exact(cuda, i, kk, eps):
    """Delayed CGS2 on a provisional basis vector ~v_i = Q_i + sum_l eps_l Q_l (normalised), with
    kk restart (Ritz) rows in the projected matrix: the step must return the corrected
    V[i] = Q_i, alpha_i = T[i, i] and the residual beta_i Q_{i+1} up to O(eps^2); using the
    uncorrected V[i] anywhere, or dropping alpha's operator term, leaves O(eps) = 1e-4."""
    from inching_amd import ops
    print("case: dcgs2 i=%d kk=%d eps=%g" % (i, kk, eps), flush=True)
    rng = np.random.default_rng(100 * i + kk)
    n = max(300, i + 100)          # i = 300: the coefficients past 256 come from global memory
    Q, A, alpha, beta, theta, betak, T = projected_system(n, i, kk, rng)
    vt = Q[:, i] + Q[:, :i] @ (eps * rng.standard_normal(i))
    vt /= np.linalg.norm(vt)
    V = np.concatenate([Q[:, :i].T, vt[None, :], np.zeros((1, n))])
    Vd, ud = _dev(V), _dev(A @ vt)
    al = _dev(np.where(np.arange(i + 2) < i, alpha[:i + 2], np.nan))      # alpha_i is the output
    bout = torch.empty(1, dtype=torch.float64, device="cuda")
    ops.dcgs2_step(Vd, i + 1, ud, al, _dev(beta), bout, vnext=Vd[i + 1])
    bar = 2e-6 if eps else 1e-12           # correct: <= 5e-7 at eps = 1e-4; each mistake: >= 6e-5
    Vo = Vd.cpu().numpy()
    assert np.array_equal(Vo[:i], Q[:, :i].T)                           # V[:i] untouched
    assert np.linalg.norm(Vo[i] - Q[:, i]) < bar
    assert abs(al[i].item() - T[i, i]) < bar
    assert abs(bout.item() - beta[i]) < bar
    assert np.linalg.norm(ud.cpu().numpy() - beta[i] * Q[:, i + 1]) < bar
    assert np.linalg.norm(Vo[i + 1] - Q[:, i + 1]) < bar


@pytest.mark.parametrize("kind", ["rcm", "random"])
def test_csr_permute_equals_fresh_search(glob2k, kind):
    """The search's CSR renumbered by a permutation equals, bit for bit, a fresh neighbour
    search of the permuted coordinates (RCM order and a random order); build_hessian's
    pattern in RCM order equals a fresh search too."""
    from inching_amd import ops
    from inching_amd.hessian import build_hessian
    print("case: csr permute %s" % kind, flush=True)
    xyz = _dev(glob2k)
    ip, ix = ops.neighbours(xyz, 8.0)
    n = xyz.shape[0]
    if kind == "rcm":
        order, undo = ops.rcm_device(ip, ix)
    else:
        perm = np.random.default_rng(5).permutation(n)
        order = _dev(perm.astype(np.int32))
        undo = _dev(np.argsort(perm).astype(np.int64))
    pp, px = ops.csr_permute(ip, ix, order, undo)
    fp, fx = ops.neighbours(xyz[order.long()].contiguous(), 8.0)
    assert torch.equal(pp, fp) and torch.equal(px, fx)
    if kind == "rcm":
        A = build_hessian(glob2k, 8.0, plus_i=1.0, reorder=True, precision=())
        gp, gx = ops.neighbours(A.xyz, 8.0)
        assert torch.equal(A.rowptr, gp) and torch.equal(A.col, gx)


def test_spmv_col16_bitwise_equal(mat):
    """16-bit column offsets (row base + offset) give bitwise the int32-column result for every
    lane count and dtype pair; a row spanning more than 16 bits is rejected."""
    from inching_amd import ops
    from inching_amd._lib import InchingError
    A, Bs = mat
    _print("case: bsr col16")
    rb, c16 = ops.bsr_col16(A.n_atoms, A.rowptr, A.col)
    rp = A.rowptr.cpu().numpy()
    col = A.col.cpu().numpy().astype(np.int64)
    rows = np.repeat(np.arange(A.n_atoms), np.diff(rp))
    assert np.array_equal(rb.cpu().numpy()[rows] + c16.cpu().numpy().astype(np.int64), col)
    rng = np.random.default_rng(4)
    x, z, acc0 = (rng.standard_normal(A.shape[0]) for _ in range(3))
    for vals, prec in ((A.vals32, torch.float32), (A.vals32, torch.float64), (A.vals64, torch.float64)):
        for lanes in (4, 8, 16, 32, 64):
            outs = []
            for c, la in ((None, lanes), ((rb, c16), lanes), ((rb, c16), lanes + 1000)):   # +1000: XCD rows
                y = torch.empty(A.shape[0], dtype=prec, device="cuda")
                acc = _dev(acc0, prec)
                ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, vals, _dev(x, prec), y, s1=0.3, c0=2.0, z=_dev(z, prec),
                             c2=1.0, acc=acc, mu=0.5, lanes=la, col16=c)
                outs.append((y.cpu().numpy(), acc.cpu().numpy()))
            for o in outs[1:]:
                assert np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1]), (lanes, prec)
    # a row whose columns span more than 65,535
    ip = torch.tensor([0, 2, 3], dtype=torch.int64, device="cuda")
    ix = torch.tensor([0, 70000, 1], dtype=torch.int32, device="cuda")
    with pytest.raises(InchingError):
        ops.bsr_col16(2, ip, ix)


def test_matfree_launch_configs_agree(mat):
    from inching_amd import ops
    from inching_amd.autotune import MATFREE_CONFIGS
    A, Bs = mat
    rng = np.random.default_rng(9)
    x, z, acc0 = (rng.standard_normal(A.shape[0]) for _ in range(3))
    ref = None
    for cfg in MATFREE_CONFIGS:
        y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
        acc = _dev(acc0)
        ops.matfree_spmv(A.xyz, A.rowptr, A.col, A.plus_i, _dev(x), y, s1=0.3, c0=2.0, z=_dev(z), c2=1.0, acc=acc,
                         mu=0.5, config=cfg)
        out = 0.3 * (Bs @ x - 2.0 * x) - z
        assert np.allclose(y.cpu().numpy(), out, rtol=1e-12, atol=1e-12), cfg
        assert np.allclose(acc.cpu().numpy(), acc0 + 0.5 * out, rtol=1e-12, atol=1e-12)
        # 16-bit column offsets (+1000), XCD-contiguous rows (+2000), both (+3000): bitwise the same
        for flag in (1000, 2000, 3000):
            y16 = torch.empty_like(y)
            acc16 = _dev(acc0)
            A.matfree_spmv(_dev(x), y16, config=flag + cfg, s1=0.3, c0=2.0, z=_dev(z), c2=1.0, acc=acc16, mu=0.5)
            assert torch.equal(y16, y) and torch.equal(acc16, acc), (flag, cfg)
        if ref is None:
            ref = y.cpu().numpy()


def _decode_clusters(cl):
    """Host decode of the cluster layout: (cluster atoms [ncl, 4], list of (j array, mask array))."""
    atoms, ptr, base, entries = (t.cpu().numpy() for t in cl)
    w = entries.view(np.uint32)
    out = []
    for c in range(base.size):
        e = w[ptr[c]:ptr[c + 1]]
        out.append((base[c] + (e >> 8).astype(np.int64), (e & 0xFF).astype(np.int64)))
    return atoms.reshape(-1, 4), out


def test_cluster_layout_reproduces_csr(mat):
    """Every neighbour pair (i, j != i) of the CSR appears exactly once in the cluster lists,
    and nothing else does (bit-exact structure); clusters partition the atoms; deterministic."""
    from inching_amd import ops
    A, _ = mat
    _print("case: cluster layout n=%d" % A.n_atoms)
    cl = A.clusters()
    assert cl is not None
    catoms, lists = _decode_clusters(cl)
    n = A.n_atoms
    flat = catoms.reshape(-1)
    assert np.array_equal(np.sort(flat[flat >= 0]), np.arange(n))
    assert np.all(flat[:n] >= 0) and np.all(flat[n:] == -1)
    pairs = []
    for c, (j, m) in enumerate(lists):
        assert np.all(np.diff(j) > 0) and np.all(m > 0) and np.all(m < 16)
        if j.size:
            assert j[0] == cl[2][c].item()
        for t in range(4):
            sel = (m >> t) & 1 == 1
            if sel.any():
                assert catoms[c, t] >= 0
                pairs.append(catoms[c, t] * np.int64(n) + j[sel])
    got = np.sort(np.concatenate(pairs))
    rp, col = A.rowptr.cpu().numpy(), A.col.cpu().numpy().astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = rows != col
    ref = np.sort(rows[off] * np.int64(n) + col[off])
    assert np.array_equal(got, ref)
    again = ops.anm_clusters(A.xyz, A.rowptr, A.col, 0.5 * A.rc, 0.25 * A.rc)
    two_calls = ops.anm_clusters(A.xyz, A.rowptr, A.col, 0.5 * A.rc, 0.25 * A.rc, one_pass=False)
    for a, b in zip(again, two_calls):          # one pass (provisional lists copied) = two merges
        assert torch.equal(a, b)
    for a, b in zip(cl, again):
        assert torch.equal(a, b)


def test_cluster_product_matches_oracle(mat, glob2k):
    """All cluster launch shapes (plain and XCD-contiguous) against the oracle's float64 block
    Hessian, with the Chebyshev epilogue and the rigid-body term; float32 vectors too."""
    from inching_amd.autotune import CLUSTER_CONFIGS
    from inching_amd.hessian import rigid_body_basis
    A, Bs = mat
    X = A.xyz.cpu().numpy()
    rp, col, vals = OH.block_hessian(X, 8.0, 1.0)
    Bo = OH.block_to_scipy(rp, col, vals, A.n_atoms)
    Q = rigid_body_basis(A.xyz)
    rng = np.random.default_rng(31)
    x, z, acc0 = (rng.standard_normal(A.shape[0]) for _ in range(3))
    qc = rng.standard_normal(Q.shape[0])
    Qh = Q.cpu().numpy()
    out = 0.3 * (Bo @ x - 2.0 * x) - z + qc @ Qh
    scale = np.abs(out).max()
    for flag in (4000, 6000):
        for cfg in CLUSTER_CONFIGS:
            _print("case: cluster product config %d" % (flag + cfg))
            y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
            acc = _dev(acc0)
            A.matfree_spmv(_dev(x), y, config=flag + cfg, s1=0.3, c0=2.0, z=_dev(z), c2=1.0, acc=acc, mu=0.5, Q=Q,
                           qcoef=_dev(qc))
            assert np.abs(y.cpu().numpy() - out).max() <= 1e-12 * scale, flag + cfg
            assert np.abs(acc.cpu().numpy() - (acc0 + 0.5 * out)).max() <= 1e-12 * max(1.0, scale)
    # plain product (no epilogue) and float32 vectors
    y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    A.matfree_spmv(_dev(x), y, config=6002)
    ref = Bo @ x
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    y32 = torch.empty(A.shape[0], dtype=torch.float32, device="cuda")
    A.matfree_spmv(_dev(x.astype(np.float32)), y32, config=6002)
    ref32 = Bo @ x.astype(np.float32).astype(np.float64)
    assert np.linalg.norm(y32.cpu().numpy() - ref32) <= 1e-6 * np.linalg.norm(ref32)


@pytest.mark.parametrize("case", ["globule2k", "padding_isolated"])
def test_persistent_cluster_product_bitwise(mat, case):
    """The persistent cluster kernel (4, 8 or 16 clusters per wave, plain and XCD-contiguous)
    against the per-cluster kernel: bitwise-equal outputs and accumulators with every
    epilogue term, float64 and float32 vectors; a globule of 2000 atoms, and 1003 atoms plus
    8 isolated ones (a padding slot, clusters with empty lists, a partial last wave)."""
    from inching_amd.hessian import build_hessian, rigid_body_basis
    if case == "globule2k":
        A, _ = mat
    else:
        X = small_globule(1003, seed=23)
        far = np.stack([np.full(8, 500.0), np.full(8, 500.0), 60.0 * np.arange(8)], axis=1)
        A = build_hessian(np.concatenate([X, far]), 8.0, plus_i=1.0, reorder=True, precision=())
        assert A.clusters() is not None
        ptr = A.clusters()[1].cpu().numpy()
        assert (np.diff(ptr) == 0).any()          # some clusters hold only isolated atoms
        assert (A.clusters()[0] == -1).any()      # and one a padding slot (1011 atoms)
    nd = A.shape[0]
    _print("case: persistent cluster product, %s, n=%d" % (case, A.n_atoms))
    Q = rigid_body_basis(A.xyz)
    rng = np.random.default_rng(77)
    x, z, acc0 = (rng.standard_normal(nd) for _ in range(3))
    qc = rng.standard_normal(Q.shape[0])
    for dt in (torch.float64, torch.float32):
        epis = ({}, dict(s1=0.3, c0=2.0, z=_dev(z, dt), c2=1.0, mu=0.5),
                dict(s1=0.3, c0=2.0, z=_dev(z, dt), c2=1.0, mu=0.5, Q=Q, qcoef=_dev(qc)))
        for k, epi in enumerate(epis):
            with_acc = "mu" in epi
            ref = torch.empty(nd, dtype=dt, device="cuda")
            ref_acc = _dev(acc0, dt) if with_acc else None
            A.matfree_spmv(_dev(x, dt), ref, config=4001, **(dict(epi, acc=ref_acc) if with_acc else epi))
            for cfg in (4104, 4108, 4116, 6104, 6108, 6116):
                y = torch.full((nd,), float("nan"), dtype=dt, device="cuda")
                acc = _dev(acc0, dt) if with_acc else None
                A.matfree_spmv(_dev(x, dt), y, config=cfg, **(dict(epi, acc=acc) if with_acc else epi))
                assert torch.equal(y, ref), (cfg, dt, k)
                if with_acc:
                    assert torch.equal(acc, ref_acc), (cfg, dt, k)


@pytest.mark.parametrize("case", ["globule2k", "padding_isolated"])
def test_record_layout_and_product(mat, case):
    """Record form of the cluster layout: the records hold each slot's coordinates, the lists
    name the same (neighbour, mask) pairs as the cluster lists in slot numbers, ascending;
    one recurrence step on records equals the cluster kernel's product with the same
    epilogue (float64, 1e-13 relative: only the order of the sum over neighbours differs),
    also with the padding slot and the empty lists of isolated atoms."""
    from inching_amd import ops
    from inching_amd.hessian import build_hessian
    if case == "globule2k":
        A, _ = mat
    else:
        X = small_globule(1003, seed=23)
        far = np.stack([np.full(8, 500.0), np.full(8, 500.0), 60.0 * np.arange(8)], axis=1)
        A = build_hessian(np.concatenate([X, far]), 8.0, plus_i=1.0, reorder=True, precision=())
    _print("case: record layout and product, %s, n=%d" % (case, A.n_atoms))
    cl = A.clusters()
    recs, rent = A.records()
    atoms = cl[0].cpu().numpy()
    ncl = atoms.size // 4
    slot_of = np.full(A.n_atoms, -1, np.int64)
    slot_of[atoms[atoms >= 0]] = np.nonzero(atoms >= 0)[0]
    rec = recs[0].cpu().numpy().reshape(-1, 6)
    xyz = A.xyz.cpu().numpy()
    assert np.array_equal(rec[atoms >= 0, :3], xyz[atoms[atoms >= 0]]) and not rec[atoms < 0].any()
    assert not rec[:, 3:].any()
    _, lists = _decode_clusters(cl)
    ptr = cl[1].cpu().numpy()
    w = rent.cpu().numpy().view(np.uint32).astype(np.int64)
    for c, (j, m) in enumerate(lists):
        e = w[ptr[c]:ptr[c + 1]]
        off = e & ~15                      # the record's byte offset, 48 bytes per slot
        assert np.all(off % 48 == 0) and np.all(np.diff(off) > 0)
        got = sorted(zip((off // 48).tolist(), (e & 15).tolist()))
        assert got == sorted(zip(slot_of[j].tolist(), m.tolist())), c
    # one step with every epilogue term, against the cluster kernel in the matrix's order
    nd = A.shape[0]
    rng = np.random.default_rng(12)
    x, z, a0 = (rng.standard_normal(nd) for _ in range(3))
    y_ref = torch.empty(nd, dtype=torch.float64, device="cuda")
    acc_ref = _dev(a0)
    A.matfree_spmv(_dev(x), y_ref, config=6001, s1=0.3, c0=2.0, z=_dev(z), c2=1.0, acc=acc_ref, mu=0.5)
    rin, rz, rout = recs
    acc = torch.empty(12 * ncl, dtype=torch.float64, device="cuda")
    ops.rec_pack(cl[0], _dev(a0), rout, acc, 1.0)          # acc = a0 in slot order
    ops.rec_pack(cl[0], _dev(z), rz)
    ops.rec_pack(cl[0], _dev(x), rin)
    for xcd in (True, False):
        acc_s = acc.clone()
        ops.matfree_rec(rin, rz, rout, cl[1], rent, A.plus_i, s1=0.3, c0=2.0, c2=1.0, acc=acc_s, mu=0.5, xcd=xcd)
        y = torch.full((nd,), float("nan"), dtype=torch.float64, device="cuda")
        yv = rout.view(-1, 6)[:, 3:].contiguous().view(-1)
        ops.rec_unpack(cl[0], yv, y)
        scale = float(y_ref.abs().max())
        assert float((y - y_ref).abs().max()) <= 1e-13 * scale, xcd
        acc_back = torch.full((nd,), float("nan"), dtype=torch.float64, device="cuda")
        ops.rec_unpack(cl[0], acc_s, acc_back)
        assert float((acc_back - acc_ref).abs().max()) <= 1e-13 * max(1.0, scale), xcd
        # padding slots stay zero
        pad = torch.as_tensor(np.repeat(atoms < 0, 3), device="cuda")
        assert not bool(yv[pad].any()) and not bool(acc_s[pad].any() if pad.any() else False)


def test_record_filter_matches_matrix_order_filter(mat):
    """A filtered application p(A) v with the recurrence on records equals the same filter
    with the vectors in the matrix's order (float64, 1e-12 relative), also in place (u = v)."""
    from inching_amd import chebyshev as CH
    from inching_amd.operator import Operator
    A, _ = mat
    _print("case: record filter vs matrix-order filter")
    lo, hi = CH.spectrum_bound(A, precision="mf64", stop_rel=CH.BOUND_STOP_REL)
    pol = CH.design_filter(A.plus_i, A.plus_i + 0.1 * (hi - A.plus_i), (A.plus_i, hi * 1.001), 5120, 5, 0.7)
    op_r = Operator(A, "mf64", pol=pol, records=True)
    op_m = Operator(A, "mf64", pol=pol, records=False)
    assert op_r.rec is not None and op_m.rec is None
    v = torch.randn(A.shape[0], dtype=torch.float64, device="cuda")
    ur, um = torch.empty_like(v), torch.empty_like(v)
    op_r.apply(v, ur)
    op_m.apply(v, um)
    s = float(um.abs().max())
    assert float((ur - um).abs().max()) <= 1e-12 * s, (float((ur - um).abs().max()), s, pol.AdjustedDegree)
    w = v.clone()
    op_r.apply(w, w)
    assert float((w - um).abs().max()) <= 1e-12 * s


def test_cluster_capacity_falls_back():
    """A cluster whose rows hold more than 4096 entries makes the layout unavailable (a host
    error, no kernel fault); the operator then keeps the per-row kernels."""
    from inching_amd import ops
    from inching_amd.autotune import best_config
    from inching_amd.hessian import build_hessian
    _print("case: cluster capacity, 1500 atoms within one cutoff")
    rng = np.random.default_rng(5)
    X = rng.uniform(0.0, 12.0, size=(1500, 3))
    A = build_hessian(X, 30.0, plus_i=1.0, reorder=True, precision=())
    with pytest.raises(ops.InchingError):
        ops.anm_clusters(A.xyz, A.rowptr, A.col, 0.5 * A.rc, 0.25 * A.rc)
    assert A.clusters() is None
    assert best_config(A, "mf64") // 1000 & 4 == 0


@pytest.mark.parametrize("nrhs", [1, 3, 7, 64])
def test_cluster_apply_multi_rhs(mat, nrhs):
    """Several right-hand sides on the cluster layout against the oracle's block Hessian and
    the per-row multi-RHS kernel."""
    from inching_amd import ops
    A, _ = mat
    _print("case: cluster apply nrhs=%d" % nrhs)
    X = A.xyz.cpu().numpy()
    rp, col, vals = OH.block_hessian(X, 8.0, 1.0)
    Bo = OH.block_to_scipy(rp, col, vals, A.n_atoms)
    rng = np.random.default_rng(40 + nrhs)
    V = rng.standard_normal((A.shape[0], nrhs))
    Y = torch.empty((A.shape[0], nrhs), dtype=torch.float64, device="cuda")
    ops.matfree_cl_apply(A.xyz, A.clusters(), A.plus_i, _dev(V), Y)
    ref = Bo @ V
    assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    Yr = torch.empty_like(Y)
    ops.matfree_apply(A.xyz, A.rowptr, A.col, A.plus_i, _dev(V), Yr)
    assert torch.allclose(Y, Yr, rtol=0, atol=1e-12 * float(np.abs(ref).max()))


def test_cluster_layout_with_padding():
    """n not a multiple of 4: the last cluster has padding slots (-1), which must neither
    read nor write; layout and products still match the CSR and the oracle."""
    from inching_amd import ops
    from inching_amd.hessian import build_hessian
    X = small_globule(1003, seed=23)
    _print("case: cluster layout with padding, n=1003")
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    cl = A.clusters()
    assert cl is not None
    catoms, lists = _decode_clusters(cl)
    flat = catoms.reshape(-1)
    assert flat.size == 1004 and np.sum(flat == -1) == 1
    assert np.array_equal(np.sort(flat[flat >= 0]), np.arange(1003))
    pairs = []
    for c, (j, m) in enumerate(lists):
        for t in range(4):
            sel = (m >> t) & 1 == 1
            if sel.any():
                assert catoms[c, t] >= 0
                pairs.append(catoms[c, t] * np.int64(1003) + j[sel])
    rp, col = A.rowptr.cpu().numpy(), A.col.cpu().numpy().astype(np.int64)
    rows = np.repeat(np.arange(1003), np.diff(rp))
    off = rows != col
    assert np.array_equal(np.sort(np.concatenate(pairs)), np.sort(rows[off] * np.int64(1003) + col[off]))
    Xr = A.xyz.cpu().numpy()
    rpo, colo, vals = OH.block_hessian(Xr, 8.0, 1.0)
    Bo = OH.block_to_scipy(rpo, colo, vals, 1003)
    rng = np.random.default_rng(3)
    x, z = rng.standard_normal(3009), rng.standard_normal(3009)
    out = 0.7 * (Bo @ x - 1.5 * x) - z
    for cfg in (4001, 6001, 6002, 6004):
        y = torch.full((3009,), np.nan, dtype=torch.float64, device="cuda")
        A.matfree_spmv(_dev(x), y, config=cfg, s1=0.7, c0=1.5, z=_dev(z), c2=1.0)
        assert np.abs(y.cpu().numpy() - out).max() <= 1e-12 * np.abs(out).max(), cfg
    V = rng.standard_normal((3009, 5))
    Y = torch.full((3009, 5), np.nan, dtype=torch.float64, device="cuda")
    ops.matfree_cl_apply(A.xyz, cl, A.plus_i, _dev(V), Y)
    ref = Bo @ V
    assert np.abs(Y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
