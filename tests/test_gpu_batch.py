"""Batched lockstep TRLM over many structures against per-structure dense eigh (oracle)."""
import numpy as np
import pytest

from oracle import hessian as OH
from tests.helpers import projected_system, small_globule

pytestmark = pytest.mark.gpu


def test_batched_solve_matches_dense(cuda):
    from inching_amd.batch import solve_batch
    sizes = [420, 700, 300, 1100, 520, 650]
    structs = [small_globule(n, seed=60 + i) for i, n in enumerate(sizes)]
    k = 8
    print("case: batch of %d structures k=%d" % (len(sizes), k), flush=True)
    r = solve_batch(structs, k=k, tol_rel=1e-10, maxiter=400)
    assert r.converged == len(sizes)
    for s, X in enumerate(structs):
        w = np.linalg.eigvalsh(OH.hessian_dense(X, 8.0, 1.0))
        assert np.allclose(r.eigval[s], w[6:6 + k], atol=1e-8), (s, r.eigval[s] - w[6:6 + k])
    assert r.rel_res.max() < 1e-9
    assert np.all(r.latency > 0) and np.all(r.latency <= r.seconds + 1e-9)


def test_batched_segmented_kernels(cuda):
    import torch
    from inching_amd.batch import SegOps, Segments
    rng = np.random.default_rng(3)
    seg_ptr = np.array([0, 5000, 5003, 16000, 30001])
    n = seg_ptr[-1]
    seg = Segments(seg_ptr, torch.device("cuda"), chunk=1024)
    so = SegOps(seg, 16, torch.device("cuda"))
    J = 11
    V = rng.standard_normal((J, n))
    u = rng.standard_normal(n)
    Vd, ud = torch.as_tensor(V, device="cuda"), torch.as_tensor(u, device="cuda")
    h = torch.empty(4 * J, dtype=torch.float64, device="cuda")
    so.multidot(Vd, J, ud, h)
    ref = np.stack([V[:, seg_ptr[s]:seg_ptr[s + 1]] @ u[seg_ptr[s]:seg_ptr[s + 1]] for s in range(4)])
    assert np.allclose(h.cpu().numpy().reshape(4, J), ref, rtol=1e-12, atol=1e-10)
    nrm = torch.empty(4, dtype=torch.float64, device="cuda")
    so.update(Vd, J, h, J, 0.5, ud, nrm=nrm)
    uref = u.copy()
    for s in range(4):
        sl = slice(seg_ptr[s], seg_ptr[s + 1])
        uref[sl] -= 0.5 * V[:, sl].T @ ref[s]
    assert np.allclose(ud.cpu().numpy(), uref, rtol=1e-11, atol=1e-10)
    assert np.allclose(nrm.cpu().numpy(), [np.linalg.norm(uref[seg_ptr[s]:seg_ptr[s + 1]]) for s in range(4)])
    W = rng.standard_normal((4, J, 5))
    out = torch.empty((5, n), dtype=torch.float64, device="cuda")
    so.combine(Vd, J, torch.as_tensor(W, device="cuda"), out)
    for s in range(4):
        sl = slice(seg_ptr[s], seg_ptr[s + 1])
        assert np.allclose(out.cpu().numpy()[:, sl], W[s].T @ V[:, sl], rtol=1e-11, atol=1e-10)
    for kk in (17, 40, 64):            # two and four 16-wide output tiles
        W = rng.standard_normal((4, J, kk))
        out = torch.empty((kk, n), dtype=torch.float64, device="cuda")
        so.combine(Vd, J, torch.as_tensor(W, device="cuda"), out)
        for s in range(4):
            sl = slice(seg_ptr[s], seg_ptr[s + 1])
            assert np.allclose(out.cpu().numpy()[:, sl], W[s].T @ V[:, sl], rtol=1e-11, atol=1e-10), (kk, s)


def test_rigid_projection_from_coordinates(cuda):
    """SegOps.rigid (the per-segment rigid-body projection from the coordinates) against the
    projection onto the QR basis of each segment's six raw rigid-body fields (numpy) and the
    stored orthonormal rows of segmented_rigid_basis: y -= scale P x out of place and in
    place, chunks that split atoms, segments far from the origin and a 3-atom segment; rigid
    motions are removed exactly by I - P."""
    import torch
    from inching_amd.batch import RigidBody, SegOps, Segments, place_apart, segmented_rigid_basis
    rng = np.random.default_rng(8)
    structs = [small_globule(700, seed=81), rng.standard_normal((3, 3)) * 3.0, small_globule(2000, seed=82),
               small_globule(1500, seed=83)]
    X, seg_atoms = place_apart(structs)
    X = X + np.array([500.0, -300.0, 800.0])
    dev = torch.device("cuda")
    xyz = torch.as_tensor(X, device=dev)
    S = len(structs)
    seg = Segments(3 * seg_atoms, dev, chunk=1024)             # 1024 dofs: chunks split atoms
    so = SegOps(seg, 8, dev)
    rig = RigidBody(xyz, seg_atoms, so)
    n3 = 3 * X.shape[0]
    x, y0 = rng.standard_normal(n3), rng.standard_normal(n3)
    ref = y0.copy()
    for s in range(S):
        a0, a1 = seg_atoms[s], seg_atoms[s + 1]
        p = X[a0:a1]
        B = np.zeros((6, a1 - a0, 3))
        for c in range(3):
            B[c, :, c] = 1.0
            e = np.zeros(3)
            e[c] = 1.0
            B[3 + c] = np.cross(e, p)
        Qs, _ = np.linalg.qr(B.reshape(6, -1).T)
        xs = x[3 * a0:3 * a1]
        ref[3 * a0:3 * a1] -= 0.7 * (Qs @ (Qs.T @ xs))
    yd = torch.as_tensor(y0, device=dev)
    so.rigid(rig, torch.as_tensor(x, device=dev), yd, 0.7)
    np.testing.assert_allclose(yd.cpu().numpy(), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    Q = segmented_rigid_basis(xyz, seg_atoms, so)
    hq = torch.empty(S * 6, dtype=torch.float64, device=dev)
    yq = torch.as_tensor(y0, device=dev)
    so.multidot(Q, 6, torch.as_tensor(x, device=dev), hq)
    so.update(Q, 6, hq, 6, 0.7, yq)
    np.testing.assert_allclose(yd.cpu().numpy(), yq.cpu().numpy(), rtol=0, atol=1e-11 * np.abs(ref).max())
    # in place: (I - P) x; a rigid motion of every segment is removed
    xi = torch.as_tensor(x, device=dev)
    so.rigid(rig, xi, xi, 1.0)
    xq = torch.as_tensor(x, device=dev)
    so.multidot(Q, 6, xq, hq)
    so.update(Q, 6, hq, 6, 1.0, xq)
    np.testing.assert_allclose(xi.cpu().numpy(), xq.cpu().numpy(), rtol=0, atol=1e-11 * np.abs(x).max())
    motion = np.zeros((X.shape[0], 3))
    for s in range(S):
        a0, a1 = seg_atoms[s], seg_atoms[s + 1]
        t, w = rng.standard_normal(3), rng.standard_normal(3) * 0.01
        motion[a0:a1] = t + np.cross(w, X[a0:a1])
    m = torch.as_tensor(motion.reshape(-1), device=dev)
    so.rigid(rig, m, m, 1.0)
    assert np.abs(m.cpu().numpy()).max() < 1e-10 * np.abs(motion).max()


@pytest.mark.parametrize("J", [11, 70, 150])
def test_segmented_fused_cgs_pass(cuda, J):
    """u -= V h per segment, second projection and norm in one pass (three tile widths)."""
    import torch
    from inching_amd.batch import SegOps, Segments
    rng = np.random.default_rng(J)
    seg_ptr = np.array([0, 5000, 5003, 16000, 30001])
    n = seg_ptr[-1]
    seg = Segments(seg_ptr, torch.device("cuda"), chunk=1024)
    so = SegOps(seg, J + 1, torch.device("cuda"))
    V = rng.standard_normal((J, n))
    u = rng.standard_normal(n)
    hh = rng.standard_normal((4, J))
    Vd, ud = torch.as_tensor(V, device="cuda"), torch.as_tensor(u, device="cuda")
    h2 = torch.empty(4 * (J + 1), dtype=torch.float64, device="cuda")
    nrm = torch.empty(4, dtype=torch.float64, device="cuda")
    so.cgs_b(Vd, J, torch.as_tensor(hh.reshape(-1), device="cuda"), J, ud, h2, nrm)
    uref = u.copy()
    for s in range(4):
        sl = slice(seg_ptr[s], seg_ptr[s + 1])
        uref[sl] -= V[:, sl].T @ hh[s]
    assert np.allclose(ud.cpu().numpy(), uref, rtol=1e-11, atol=1e-10)
    got = h2.cpu().numpy().reshape(4, J + 1)
    for s in range(4):
        sl = slice(seg_ptr[s], seg_ptr[s + 1])
        assert np.allclose(got[s, :J], V[:, sl] @ uref[sl], rtol=1e-10, atol=1e-8)
        assert np.isclose(got[s, J], uref[sl] @ uref[sl], rtol=1e-12)
        assert np.isclose(nrm.cpu().numpy()[s], np.linalg.norm(uref[sl]), rtol=1e-12)


@pytest.mark.parametrize("i", [0, 1, 19, 59])
def test_segmented_dcgs2_step_first_order_exact(cuda, i):
    """Segmented delayed CGS2: every segment holds its own basis with a known projected matrix
    (arrowhead rows for some) and a provisional vector perturbed by 1e-4; per segment the
    corrected V[i], alpha_i, beta_i and the residual must agree to < 1e-6, where a wrong
    or missing correction leaves O(1e-4). Segments are cut into several chunks."""
    import torch
    from inching_amd.batch import SegOps, Segments
    print("case: seg dcgs2 i=%d" % i, flush=True)
    dofs, kks = [300, 257, 411, 1000], [0, 0, min(6, max(i - 1, 0)), max(i - 2, 0)]
    seg_ptr = np.concatenate([[0], np.cumsum(dofs)])
    n, S = int(seg_ptr[-1]), len(dofs)
    seg = Segments(seg_ptr, torch.device("cuda"), chunk=128)
    so = SegOps(seg, i + 2, torch.device("cuda"))
    V = np.zeros((i + 2, n))
    u = np.zeros(n)
    want = []
    alpha = np.full((i + 2, S), np.nan)
    beta = np.zeros((i + 2, S))
    for s, (m, kk) in enumerate(zip(dofs, kks)):
        rng = np.random.default_rng(1000 * i + s)
        Q, A, al, be, _, _, T = projected_system(m, i, kk, rng)
        vt = Q[:, i] + Q[:, :i] @ (1e-4 * rng.standard_normal(i))
        vt /= np.linalg.norm(vt)
        sl = slice(seg_ptr[s], seg_ptr[s + 1])
        V[:i, sl] = Q[:, :i].T
        V[i, sl] = vt
        u[sl] = A @ vt
        alpha[:i, s] = al[:i]
        beta[:, s] = be[:i + 2]
        want.append((sl, Q, T[i, i], be[i]))
    Vd, ud = torch.as_tensor(V, device="cuda"), torch.as_tensor(u, device="cuda")
    ad, bd = torch.as_tensor(alpha, device="cuda"), torch.as_tensor(beta, device="cuda")
    bout = torch.empty(S, dtype=torch.float64, device="cuda")
    so.dcgs2(Vd, i + 1, ud, ad[i], bd[i - 1] if i > 0 else None, bout)
    Vo, uo, ao, bo = Vd.cpu().numpy(), ud.cpu().numpy(), ad.cpu().numpy(), bout.cpu().numpy()
    assert np.array_equal(Vo[:i], V[:i])
    for s, (sl, Q, a_ref, b_ref) in enumerate(want):
        assert np.linalg.norm(Vo[i, sl] - Q[:, i]) < 2e-6, s
        assert abs(ao[i, s] - a_ref) < 2e-6, s
        assert abs(bo[s] - b_ref) < 2e-6, s
        assert np.linalg.norm(uo[sl] - b_ref * Q[:, i + 1]) < 2e-6, s


def test_batched_solve_dcgs2_matches_cgs2(cuda):
    """The lockstep solve with delayed CGS2 and with fused CGS2 agree on every eigenvalue."""
    from inching_amd.batch import solve_batch
    print("case: batch dcgs2 vs cgs2", flush=True)
    structs = [small_globule(n, seed=70 + i) for i, n in enumerate([380, 520, 450])]
    r1 = solve_batch(structs, k=8, tol_rel=1e-10, maxiter=400)
    r2 = solve_batch(structs, k=8, tol_rel=1e-10, maxiter=400, dcgs2=False)
    assert r1.converged == r2.converged == len(structs)
    assert np.allclose(r1.eigval, r2.eigval, rtol=0, atol=1e-9)
    assert r1.rel_res.max() < 1e-9


@pytest.mark.parametrize("filtered,per_component", [(False, 12), (True, 12), (True, 4)])
def test_assembly_merge_matches_union_of_component_spectra(cuda, filtered, per_component):
    """Disconnected assembly: the lowest modes are the merged per-subunit spectra (the
    oracle: dense eigh of every subunit Hessian; 6 rigid modes per subunit excluded). With
    4 modes per component (20 < 30) the coverage check sends the solve round again with more
    modes per component."""
    from inching_amd.batch import solve_assembly
    from inching_amd.synth import icosahedral_group
    print("case: assembly merge filtered=%s per_component=%d" % (filtered, per_component), flush=True)
    G = icosahedral_group()[:5]
    subs = [small_globule(400 + 40 * i, seed=90 + i) for i in range(5)]
    u = np.array([0.3, 0.5, 0.81])
    u /= np.linalg.norm(u)
    parts = [s @ G[i].T + 200.0 * (G[i] @ u) for i, s in enumerate(subs)]
    X = np.concatenate(parts)
    k_total = 30
    r = solve_assembly(X, k_total, tol_rel=1e-9, per_component=per_component, filtered=filtered, max_degree=2000)
    assert r.components == 5
    if per_component == 4:
        assert r.per_component > 4
    ref = np.sort(np.concatenate([np.linalg.eigvalsh(OH.hessian_dense(p, 8.0, 1.0))[6:] for p in parts]))[:k_total]
    assert np.allclose(r.eigval, ref, atol=1e-8), np.max(np.abs(r.eigval - ref))
    assert r.rel_res.max() < 1e-8
