"""The reference's own workflow, written against the drop-in modules, checked with the oracle.

Mirrors the driver protocol of ``Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py``
and the notebooks: coordinates in nm -> 3DRCM -> islands + HalfNnz -> Hessian (A, A_diag) ->
rigid-body Q -> solver -> residual ||A v + A^T v - A_diag v - lam v|| per mode.
"""
import numpy as np
import pytest
import torch

from oracle import chebyshev as OC
from oracle import geometry as OG
from oracle import hessian as OH
from oracle import partition as OP
from oracle import rcm as OR
from tests.helpers import small_globule, subspace_sines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workflow(cuda):
    from InchingLiteInteger.Fuel.Coordinate.T1 import X_KdCuthillMckeeOrder, X_KdUngappedMinMaxNeighbor
    from InchingLiteInteger.Fuel.T1 import Xnumpy_SparseCupyMatrixUngappped
    X = small_globule(900, seed=41) / 10.0        # nm, as the reference expects
    order, undo = X_KdCuthillMckeeOrder(X, rc_Gamma=8.0, Reverse=True)
    Xp = X[order]
    islands, half = X_KdUngappedMinMaxNeighbor(Xp, rc_Gamma=8.0, maxleafsize=100, User_ReturnHalfNnz=True,
                                               SliceForm=True)
    builder = Xnumpy_SparseCupyMatrixUngappped(Xp, batch_head=None, maxleafsize=100, rc_Gamma=8.0, User_PlusI=1.0,
                                               NnzMinMaxDict=islands)
    A, A_diag = builder.ReturnCupyHLowerTriangleInt32(User_MaxHalfNnzBufferSize=half)
    Q = OH.rigid_body_basis(Xp)               # the driver's construction, numpy
    Ad = OH.hessian_dense(Xp, 0.8, 1.0)
    w, U = np.linalg.eigh(Ad)
    return dict(X=X, order=order, undo=undo, Xp=Xp, islands=islands, half=half, builder=builder, A=A,
                A_diag=A_diag, Q=Q, Ad=Ad, w=w, U=U)


def test_rcm_and_islands_match_oracle(workflow):
    print("case: shim rcm + islands", flush=True)
    ip, ix = OG.neighbour_csr(workflow["X"], 0.8)
    o_ref, u_ref = OR.rcm_order(ip, ix)
    assert np.array_equal(workflow["order"], o_ref) and np.array_equal(workflow["undo"], u_ref)
    ip2, ix2 = OG.neighbour_csr(workflow["Xp"], 0.8)
    isl_ref, half_ref = OP.ungapped_islands(ip2, ix2, maxleafsize=100, gap_size=100)
    assert workflow["islands"] == isl_ref and workflow["half"] == half_ref
    assert workflow["builder"].ReturnNumberTotalBatch() == len(OP.batch_head(900, 100))
    # the reference returns the slice-form islands whatever SliceForm says (Fuel/Coordinate/T1.py:470-481)
    from InchingLiteInteger.Fuel.Coordinate.T1 import X_KdUngappedMinMaxNeighbor
    isl_f = X_KdUngappedMinMaxNeighbor(workflow["Xp"], rc_Gamma=8.0, maxleafsize=100, SliceForm=False)
    assert isl_f == isl_ref


def test_lower_triangle_semantics(workflow):
    A, A_diag = workflow["A"], workflow["A_diag"]
    L, d = OH.lower_csr_and_diag(workflow["Xp"], 0.8, 1.0)
    v = np.random.default_rng(0).standard_normal(A.shape[0])
    vd = torch.as_tensor(v, device="cuda")
    assert np.allclose((A @ vd).cpu().numpy(), L @ v, rtol=1e-12, atol=1e-12)
    assert np.allclose((A.T @ vd).cpu().numpy(), L.T @ v, rtol=1e-12, atol=1e-12)
    assert np.allclose(A_diag.cpu().numpy(), d, rtol=1e-13)
    from InchingLiteInteger.Burn.Krylov.T3 import OOC2_HalfMemS_v_KrylovAv_VOID
    aux = OOC2_HalfMemS_v_KrylovAv_VOID(A, A_diag)
    u = torch.empty_like(vd)
    aux(A, vd, u)
    assert np.allclose(u.cpu().numpy(), OH.halfmem_matvec(L, d, v), rtol=1e-12, atol=1e-12)


def test_hotelling_and_orthogonalization(workflow):
    from InchingLiteInteger.Burn.Orthogonalization.T2 import (OOC2_qnext_Q_ICGSqnext, OOC2_qnext_Q_MGSqnext,
                                                             T2_vnext_V_MGSvnext)
    from InchingLiteInteger.Burn.Orthogonalization.T3 import T3_QHotelling_x_Ax_HotelledAx
    Q = workflow["Q"]
    n = Q.shape[1]
    rng = np.random.default_rng(1)
    x, res = rng.standard_normal(n), rng.standard_normal(n)
    rd = torch.as_tensor(res.copy(), device="cuda")
    T3_QHotelling_x_Ax_HotelledAx(Q, torch.as_tensor(x, device="cuda"), rd, HotellingShift=10.0)
    assert np.allclose(rd.cpu().numpy(), OH.hotelling(Q, x, res.copy(), 10.0), rtol=1e-12, atol=1e-12)
    V = np.linalg.qr(rng.standard_normal((n, 9)))[0]
    u = rng.standard_normal(n)
    ud = torch.as_tensor(u.copy(), device="cuda")
    T2_vnext_V_MGSvnext(ud, torch.as_tensor(V, device="cuda"))
    assert np.allclose(ud.cpu().numpy(), u - V @ (V.T @ u), atol=1e-12)
    z = OOC2_qnext_Q_ICGSqnext(torch.as_tensor(u, device="cuda"), torch.as_tensor(V, device="cuda")).cpu().numpy()
    assert abs(np.linalg.norm(z) - 1) < 1e-12 and np.max(np.abs(V.T @ z)) < 1e-12
    z2 = OOC2_qnext_Q_MGSqnext(torch.as_tensor(u, device="cuda"), torch.as_tensor(V, device="cuda")).cpu().numpy()
    assert np.max(np.abs(V.T @ z2)) < 1e-12


def _driver_residuals(A, A_diag, lam, V):
    out = []
    for j in range(V.shape[1]):
        v = V[:, j].contiguous()
        B = (A @ v) + (A.T @ v) - A_diag * v - lam[j] * v
        out.append(float(torch.linalg.norm(B)))
    return np.array(out)


def test_trlm_shim_matches_dense(workflow):
    from InchingLiteInteger.Burn.ThickRestartLanczosHotellingDeflation.T1 import S_HeigvalTRLMHD_HeigvecTRLMHD
    print("case: shim TRLM k=16", flush=True)
    A, A_diag, w, U = workflow["A"], workflow["A_diag"], workflow["w"], workflow["U"]
    k = 16
    lam, V = S_HeigvalTRLMHD_HeigvecTRLMHD(A, A_diag, k=k, tol=1e-10, User_Q_HotellingDeflation=workflow["Q"],
                                           User_HotellingShift=10.0)
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[6:6 + k], atol=1e-9)
    res = _driver_residuals(A, A_diag, lam, V)
    assert res.max() < 1e-8
    assert subspace_sines(U[:, 6:6 + k], V.cpu().numpy()).max() < 1e-4


def test_matfree_operator_on_nm_coordinates(workflow):
    """The drop-in works in nm (squared distances 0.01-0.64 nm^2 here, not the Angstrom range
    of the other operator tests): the matrix-free float64 products (the autotuned single-vector
    kernel and the multi-vector one) against the oracle's dense Hessian, and a TRLM solve with
    precision='mf64' against dense eigh."""
    from InchingLiteInteger.Burn.ThickRestartLanczosHotellingDeflation.T1 import S_HeigvalTRLMHD_HeigvecTRLMHD
    from inching_amd.operator import Operator
    print("case: shim mf64 on nm coordinates", flush=True)
    A, Ad, w = workflow["A"], workflow["Ad"], workflow["w"]
    rng = np.random.default_rng(5)
    V = rng.standard_normal((Ad.shape[0], 3))
    Vd = torch.as_tensor(V, device="cuda")
    ref = Ad @ V
    got = A.matfree(Vd.contiguous()).cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    op = Operator(A, "mf64")
    u = torch.empty(Ad.shape[0], dtype=torch.float64, device="cuda")
    op.plain(Vd[:, 0].contiguous(), u)
    assert np.max(np.abs(u.cpu().numpy() - ref[:, 0])) <= 1e-13 * np.max(np.abs(ref[:, 0]))
    k = 12
    lam, Vk = S_HeigvalTRLMHD_HeigvecTRLMHD(A, workflow["A_diag"], k=k, tol=1e-10,
                                            User_Q_HotellingDeflation=workflow["Q"], User_HotellingShift=10.0,
                                            precision="mf64")
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[6:6 + k], atol=1e-9)
    assert _driver_residuals(A, workflow["A_diag"], lam, Vk).max() < 1e-8


def test_ctrlm_protocol_shim(workflow):
    """Spectrum bound -> KPM-optimised filter -> filtered TRLM (the v023 notebook protocol)."""
    from InchingLiteInteger.Burn.HermitianLanczos.T2 import A_Adiag_EstimateSpectrumBound
    from InchingLiteInteger.Burn.PolynomialFilters.T2 import A_Adiag_OptimizePolynomialParamsOnMemory
    from InchingLiteInteger.Burn.ThickRestartLanczosHotellingDeflation.T1 import S_HeigvalTRLMHD_HeigvecTRLMHD
    print("case: shim CTRLM k=10", flush=True)
    A, A_diag, w = workflow["A"], workflow["A_diag"], workflow["w"]
    k = 10
    sb = A_Adiag_EstimateSpectrumBound(A, A_diag)
    assert sb[1].get() <= w[-1] * (1 + 1e-9) and sb[1] >= w[-1] * (1 - 1e-6)
    bound = (1.0, sb[1].get() + 1e-12)
    pol, cnt, interval = A_Adiag_OptimizePolynomialParamsOnMemory(
        A, A_diag, User_MaximumDegree=5120, User_MinimumDegree=5, User_DampingKernel="Jackson",
        User_ExtremalIntervalDefinition=1e-10, User_WantedInterval=(1.0, bound[1] / 10), User_SpectrumBound=bound,
        User_DesignatedStart=bound[1] / 10, User_WantedNumberEigenvalue=(k - 5) * 3 + 20,
        User_AffordableMemoryMargin=25, User_NumberKpmTrials=5, User_ConvergenceRatio=0.7)
    # the chosen filter must match the reference design for the interval it settled on
    ref = OC.optimize_on_degree(5120, 5, "Jackson", 1e-10, interval, bound, 0.7)
    assert pol.AdjustedDegree == ref.AdjustedDegree
    assert np.array_equal(np.asarray(pol.AdjustedCoefficients), np.asarray(ref.AdjustedCoefficients))
    lam, V = S_HeigvalTRLMHD_HeigvecTRLMHD(A, A_diag, k=k, tol=1e-9, User_Q_HotellingDeflation=workflow["Q"],
                                           User_HotellingShift=-40.0, User_PolynomialParams=pol)
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[6:6 + k], atol=1e-8)


def test_chebyshev_av_and_kpm(workflow):
    from InchingLiteInteger.Burn.PolynomialFilters.T0 import OOC0_OptimizePolynomialParamsOnDegree
    from InchingLiteInteger.Burn.PolynomialFilters.T2 import (A_Adiag_PolynomialParams_KpmEstimate,
                                                             OOC2_A_Adiag_ChebyshevAv)
    A, A_diag, w, U, Ad = workflow["A"], workflow["A_diag"], workflow["w"], workflow["U"], workflow["Ad"]
    pol = OOC0_OptimizePolynomialParamsOnDegree(2000, 5, "Jackson", 1e-10, (1.0, 3.0), (1.0, w[-1] + 1e-12), 0.7)
    v = np.random.default_rng(2).standard_normal(A.shape[0])
    C = OOC2_A_Adiag_ChebyshevAv(A, A_diag, User_PolynomialParams=pol)
    y = C.ChebyshevAv(A, torch.as_tensor(v, device="cuda")).cpu().numpy()
    y_ref = OC.cheb_av(lambda x: Ad @ x, v, pol)
    assert np.allclose(y, y_ref, rtol=1e-9, atol=1e-9)
    y2, rho = C.ChebyshevAv(A, torch.as_tensor(v / np.linalg.norm(v), device="cuda"), User_ReturnRho=True)
    _, rho_ref = OC.cheb_av(lambda x: Ad @ x, v / np.linalg.norm(v), pol, return_rho=True)
    assert abs(rho - rho_ref) < 1e-9
    cnt = A_Adiag_PolynomialParams_KpmEstimate(A, A_diag, User_PolynomialParams=pol, User_NumberKpmTrials=40)
    exact = int(np.sum((w >= 1.0 - 1e-9) & (w <= 3.0)))
    # KPM is a stochastic count (40 probes): the reference adds 1 and tolerates its spread
    assert abs(cnt - exact) <= max(25, 0.5 * exact)


def test_jdm_shim_matches_dense(workflow):
    from InchingLiteInteger.Burn.JacobiDavidsonHotellingDeflation.T1 import S_HeigvalJDMHD_HeigvecJDMHD
    print("case: shim JDM k=8", flush=True)
    A, A_diag, w = workflow["A"], workflow["A_diag"], workflow["w"]
    k = 8
    lam, V = S_HeigvalJDMHD_HeigvecJDMHD(A, A_diag, k=k, tol=1e-9, maxiter=3000, User_GapEstimate=0.0,
                                         User_Q_HotellingDeflation=workflow["Q"], User_HotellingShift=40.0)
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[6:6 + k], atol=1e-9)
    assert _driver_residuals(A, A_diag, lam, V).max() < 1e-8


def test_irlm_shim_matches_dense(workflow):
    from InchingLiteInteger.Burn.ImplicitlyRestartedLanczosHotellingDeflation.T1 import S_HeigvalIRLMHD_HeigvecIRLMHD
    print("case: shim IRLM k=32", flush=True)
    A, w = workflow["A"], workflow["w"]
    k = 32
    lam, V = S_HeigvalIRLMHD_HeigvecIRLMHD(A, k=k, tol=1e-10, User_Q_HotellingDeflation=workflow["Q"],
                                           User_HotellingShift=10.0, maxiter=3000)
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[6:6 + k], atol=1e-9), np.max(np.abs(lam - w[6:6 + k]))
    # the driver's residual per mode and the principal angles to the dense eigenvectors (the
    # 38th and 39th eigenvalues are 0.063 apart, so the 32-mode subspace is well separated)
    res = _driver_residuals(workflow["A"], workflow["A_diag"], lam, V)
    assert res.max() < 1e-7, res.max()
    Vn = V.cpu().numpy()
    assert np.max(np.abs(Vn.T @ Vn - np.eye(k))) < 1e-10
    assert subspace_sines(workflow["U"][:, 6:6 + k], Vn).max() < 1e-5


def test_cdsi_shim_matches_dense(workflow):
    """Chebyshev-Davidson (no deflation, as the reference): the lowest 16 eigenpairs including
    the six rigid-body modes at PlusI; the assertions of the reference."""
    from InchingLiteInteger.Burn.ChebyshevDavidsonSubspaceIteration.T1 import S_HeigvalCDSIHD_HeigvecCDSIHD
    from InchingLiteInteger.Burn.HermitianLanczos.T2 import A_Adiag_EstimateSpectrumBound
    print("case: shim CDSI k=16", flush=True)
    A, A_diag, w = workflow["A"], workflow["A_diag"], workflow["w"]
    bound = A_Adiag_EstimateSpectrumBound(A, A_diag)
    bound = (1.0, bound[1].get() + 1e-12)
    k = 16
    lam, V = S_HeigvalCDSIHD_HeigvecCDSIHD(A, A_diag, k=k, tol=1e-12, maxiter=20000, User_ChebyshevDegree=20,
                                           User_SpectrumBound=bound)
    lam = lam.cpu().numpy()
    assert np.allclose(lam, w[:k], atol=1e-9)
    assert _driver_residuals(A, A_diag, lam, V).max() < 1e-8
    with pytest.raises(AssertionError, match="ABORTED"):
        S_HeigvalCDSIHD_HeigvecCDSIHD(A, A_diag, k=k, User_SpectrumBound=None)
    with pytest.raises(AssertionError, match="ABORTED"):
        S_HeigvalCDSIHD_HeigvecCDSIHD(A, A_diag, k=k, User_SpectrumBound=bound,
                                       User_Q_HotellingDeflation=workflow["Q"])
