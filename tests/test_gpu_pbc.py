"""Periodic x/y images (membrane boxes): neighbour masks, image-summed Hessian blocks and
the periodic matrix-free operator against the brute-force oracle, and a periodic solve
against dense eigh (three translational zero modes)."""
import numpy as np
import pytest
import torch

from oracle import hessian as OH

pytestmark = pytest.mark.gpu


def _slab(n, box, height, seed, dmin=2.0):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform([0, 0, 0], [box[0], box[1], height])
        ok = True
        for q in pts[-400:]:
            d = np.abs(p - q)
            d[:2] = np.minimum(d[:2], np.array(box) - d[:2])
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


def _dense(rp, col, vals, n):
    D = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            D[3 * i:3 * i + 3, 3 * col[e]:3 * col[e] + 3] = vals[e]
    return D


CASES = {"narrow_box_multi_image": ((15.0, 15.0), 8.0, 90), "membrane_slab": ((40.0, 36.0), 12.0, 700)}


@pytest.mark.parametrize("case", list(CASES))
def test_pbc_pattern_assembly_and_operator(cuda, case):
    from inching_amd import ops
    box, height, n = CASES[case]
    print("case: pbc %s n=%d" % (case, n), flush=True)
    X = _slab(n, box, height, seed=3)
    xyz = torch.as_tensor(X, device="cuda")
    rp_o, col_o, v_o, m_o = OH.block_hessian_pbc(X, 8.0, 1.0, box)
    ip, ix, im = ops.neighbours_pbc(xyz, 8.0, box, (0.0, 0.0))
    assert np.array_equal(ip.cpu().numpy(), rp_o) and np.array_equal(ix.cpu().numpy(), col_o)
    assert np.array_equal(im.cpu().numpy().astype(np.int64), m_o)
    if case == "narrow_box_multi_image":
        assert np.any([bin(int(x)).count("1") > 1 for x in m_o])
    v64, v32, diag = ops.hessian_assemble_pbc(xyz, 1.0, ip, ix, im, box)
    v64 = v64.cpu().numpy().reshape(-1, 3, 3)
    scale = np.abs(v_o).max()
    assert np.max(np.abs(v64 - v_o)) <= 1e-13 * scale
    assert np.max(np.abs(v32.cpu().numpy().reshape(-1, 3, 3) - v_o)) <= 1e-6 * scale
    D = _dense(rp_o, col_o, v_o, n)
    rng = np.random.default_rng(1)
    x, z, acc0 = (rng.standard_normal(3 * n) for _ in range(3))
    y = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    acc = torch.as_tensor(acc0.copy(), device="cuda")
    ops.matfree_pbc(xyz, ip, ix, im, box, 1.0, torch.as_tensor(x, device="cuda"), y, s1=0.3, c0=2.0,
                    z=torch.as_tensor(z, device="cuda"), c2=1.0, acc=acc, mu=0.5)
    want = 0.3 * (D @ x - 2.0 * x) - z
    assert np.allclose(y.cpu().numpy(), want, rtol=1e-12, atol=1e-12 * scale)
    assert np.allclose(acc.cpu().numpy(), acc0 + 0.5 * want, rtol=1e-12, atol=1e-12 * scale)
    Xr = rng.standard_normal((3 * n, 5))
    Y = torch.empty((3 * n, 5), dtype=torch.float64, device="cuda")
    ops.matfree_pbc(xyz, ip, ix, im, box, 1.0, torch.as_tensor(Xr, device="cuda"), Y)
    assert np.allclose(Y.cpu().numpy(), D @ Xr, rtol=1e-12, atol=1e-12 * scale)


def test_pbc_solve_matches_dense(cuda):
    """RCM-ordered periodic Hessian, filtered TRLM with only the three translations deflated
    (periodic images break rotational invariance), against dense eigh."""
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes
    print("case: pbc solve", flush=True)
    box = (40.0, 36.0)
    X = _slab(700, box, 12.0, seed=5)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f64",), box=box)
    Q = rigid_body_basis(A.xyz, rotations=False)
    rp_o, col_o, v_o, _ = OH.block_hessian_pbc(X[A.order], 8.0, 1.0, box)
    w = np.linalg.eigvalsh(_dense(rp_o, col_o, v_o, 700))
    assert np.sum(np.abs(w - 1.0) < 1e-9) == 3
    k = 8
    for prec in ("mf64", "f64"):
        r = lowest_modes(A, k, Q=Q, precision=prec, tol_rel=1e-9)
        assert np.max(np.abs(r.eigval - w[3:3 + k])) <= 1e-9 * w[-1], prec
        assert np.max(r.rel_res) <= 1e-9


def test_pbc_shim_builder_and_box_assertions(cuda):
    from InchingLiteInteger.Fuel.Coordinate.T1 import X_KdCuthillMckeeOrder
    from InchingLiteInteger.Fuel.T1 import Xnumpy_SparseCupyMatrixUngappped
    print("case: pbc shim", flush=True)
    box = (40.0, 36.0)
    X = _slab(400, box, 10.0, seed=7) / 10.0                  # nm, as the reference's loaders
    pbc = {"RectBox_Xsize": 4.0, "RectBox_Ysize": 3.6, "RectBox_Zsize": 1.0, "X": (0.0, 4.0), "Y": (0.0, 3.6)}
    order, undo = X_KdCuthillMckeeOrder(X, rc_Gamma=8.0, User_DictCharmmGuiPbc=pbc)
    assert sorted(order.tolist()) == list(range(400))
    Xo = X[order]
    A, A_diag = Xnumpy_SparseCupyMatrixUngappped(Xo, rc_Gamma=8.0, User_PlusI=1.0,
                                                 User_DictCharmmGuiPbc=pbc).ReturnCupyHLowerTriangleInt32()
    rp_o, col_o, v_o, _ = OH.block_hessian_pbc(Xo, 0.8, 1.0, (4.0, 3.6))
    D = _dense(rp_o, col_o, v_o, 400)
    v = np.random.default_rng(2).standard_normal(1200)
    got = A.matvec(torch.as_tensor(v, device="cuda")).cpu().numpy()
    assert np.allclose(got, D @ v, rtol=1e-12, atol=1e-12 * np.abs(D).max())
    bad = dict(pbc, RectBox_Xsize=0.5)
    with pytest.raises(AssertionError, match="ABORTED"):
        X_KdCuthillMckeeOrder(X, rc_Gamma=8.0, User_DictCharmmGuiPbc=bad)
