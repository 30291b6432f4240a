"""Pins the CPU oracle: golden vectors, library routines, closed forms and invariants."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from oracle import chebyshev as OC
from oracle import geometry as OG
from oracle import hessian as OH
from oracle import gnn as ON
from oracle import lanczos as OL
from oracle import partition as OP
from oracle import rcm as OR
from tests.helpers import small_globule, subspace_sines

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ geometry --
def test_neighbours_kdtree_equals_bruteforce():
    rng = np.random.default_rng(1)
    X = rng.uniform(0, 25, size=(600, 3))
    a = OG.neighbour_csr(X, 6.5)
    b = OG.neighbour_csr_bruteforce(X, 6.5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # symmetric, self included
    ip, ix = a
    rows = np.repeat(np.arange(600), np.diff(ip))
    S = sp.csr_matrix((np.ones(ix.size), (rows, ix)), shape=(600, 600))
    assert (S != S.T).nnz == 0
    assert np.all(S.diagonal() == 1)


# ----------------------------------------------------------------------- RCM --
def test_rcm_path_graph_known_answer():
    # path 0-1-2-3-4 (self included): seeds by stable degree -> 0; BFS 0,1,2,3,4; reversed
    ip = np.array([0, 2, 5, 8, 11, 13])
    ix = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4])
    order, undo = OR.rcm_order(ip, ix)
    assert order.tolist() == [4, 3, 2, 1, 0]
    assert np.array_equal(undo, np.argsort(order))


def test_rcm_children_sorted_by_degree_then_descending_index():
    # star centre 0 with leaves 1..4; leaf 2 also touches 5 (higher degree)
    edges = [(0, 1), (0, 2), (0, 3), (0, 4), (2, 5)]
    n = 6
    adj = [set([i]) for i in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    ip = np.zeros(n + 1, dtype=np.int64)
    ip[1:] = np.cumsum([len(s) for s in adj])
    ix = np.concatenate([sorted(s) for s in adj])
    order, _ = OR.rcm_order(ip, ix, reverse=False)
    # seed = first min-degree node (stable): degrees [5,2,3,2,2,2] -> node 1
    # BFS from 1: neighbours desc [1,0] -> 0; from 0: desc [4,3,2,1,0] new 4,3,2 sorted by degree (2,2,3) -> 4,3,2
    # from 4: none; 3: none; 2: desc [5,2,0] -> 5
    assert order.tolist() == [1, 0, 4, 3, 2, 5]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rcm_matches_scipy_reverse_cuthill_mckee(seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 30, size=(900, 3))
    ip, ix = OG.neighbour_csr(X, 6.0)
    n = len(ip) - 1
    ixd = np.concatenate([ix[ip[i]:ip[i + 1]][::-1] for i in range(n)])
    G = sp.csr_matrix((np.ones(len(ixd)), ixd, ip), shape=(n, n))
    ref = reverse_cuthill_mckee(G, symmetric_mode=True)
    deg = np.diff(ip)
    order, _ = OR.rcm_order(ip, ix, seed_order=np.argsort(deg))
    assert np.array_equal(order, ref)


@pytest.mark.parametrize("case", ["globule", "two_blobs", "scipy_seed", "single_atom"])
def test_c_twin_matches_numpy_oracle(case):
    """The plain-C restatement (cell grid + BFS) and the numpy one (k-d tree) give the same
    neighbour CSR and the same permutation; with numpy's default argsort as the seed order the
    C twin equals scipy's reverse_cuthill_mckee too."""
    from oracle import c_twin as CT
    rng = np.random.default_rng(11)
    if case == "single_atom":
        X = np.zeros((1, 3))
    elif case == "two_blobs":
        X = np.concatenate([rng.uniform(0, 15, (300, 3)), rng.uniform(60, 70, (120, 3))])
    else:
        X = small_globule(1500, seed=5)
    a = OG.neighbour_csr(X, 8.0)
    b = CT.neighbour_csr(X, 8.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    seed_order = np.argsort(np.diff(a[0])) if case == "scipy_seed" else None
    o1, u1 = OR.rcm_order(*a, seed_order=seed_order)
    o2, u2 = CT.rcm_order(*b, seed_order=seed_order)
    assert np.array_equal(o1, o2) and np.array_equal(u1, u2)
    if case == "scipy_seed":
        n = len(a[0]) - 1
        ixd = np.concatenate([a[1][a[0][i]:a[0][i + 1]][::-1] for i in range(n)])
        G = sp.csr_matrix((np.ones(len(ixd)), ixd, a[0]), shape=(n, n))
        assert np.array_equal(o2, reverse_cuthill_mckee(G, symmetric_mode=True))


@pytest.mark.parametrize("seed", range(6))
def test_level_synchronous_rcm_equals_sequential(seed):
    """The level-synchronous restatement (owner = first adjacent node of the previous level,
    then degree, then descending index) reproduces the sequential order, on random
    geometric graphs with several components and on random non-geometric graphs."""
    rng = np.random.default_rng(seed)
    if seed % 2:
        n = 300
        rows = [np.unique(np.concatenate([[i], rng.integers(0, n, rng.integers(0, 6))])) for i in range(n)]
        A = sp.lil_matrix((n, n))
        for i, r in enumerate(rows):
            A[i, r] = 1
            A[r, i] = 1
        A = A.tocsr()
        A.sort_indices()
        ip, ix = A.indptr.astype(np.int64), A.indices.astype(np.int64)
    else:
        X = np.concatenate([rng.uniform(0, 20, (250, 3)), rng.uniform(50, 60, (80, 3)), rng.uniform(-90, -80, (3, 3))])
        ip, ix = OG.neighbour_csr(X, 5.0)
    for so in (None, np.argsort(np.diff(ip))):
        for rev in (True, False):
            a = OR.rcm_order(ip, ix, seed_order=so, reverse=rev)
            b = OR.rcm_order_levels(ip, ix, seed_order=so, reverse=rev)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _all_cuthill_mckee_orders(adj, deg):
    """Every order a Cuthill-McKee BFS can produce when ties are broken arbitrarily: a new
    component starts at an unvisited node of minimal degree, and each dequeued node appends
    its unvisited neighbours in non-decreasing degree (any order within equal degree)."""
    import itertools
    n = len(adj)
    out = set()

    def rec(order, ptr):
        if len(order) == n:
            out.add(tuple(order))
            return
        if ptr < len(order):
            node = order[ptr]
            kids = sorted({j for j in adj[node] if j not in order})
            groups = [list(g) for _, g in itertools.groupby(sorted(kids, key=lambda j: deg[j]), key=lambda j: deg[j])]
            for combo in itertools.product(*[list(itertools.permutations(g)) for g in groups]):
                rec(order + [j for g in combo for j in g], ptr + 1)
            return
        rest = [j for j in range(n) if j not in order]
        dmin = min(deg[j] for j in rest)
        for sd in rest:
            if deg[sd] == dmin:
                rec(order + [sd], ptr)

    rec([], 0)
    return out


@pytest.mark.parametrize("seed", range(8))
def test_rcm_is_a_cuthill_mckee_order_brute_force(seed):
    """On tiny random graphs the (unreversed) order is one of the enumerated Cuthill-McKee
    orders, i.e. level structure, degree sorting and seed choice are right independently of
    the tie rule."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(4, 8))
    adj = [{i} for i in range(n)]
    for _ in range(int(rng.integers(2, 9))):
        a, b = rng.integers(0, n, 2)
        adj[a].add(int(b))
        adj[b].add(int(a))
    ip = np.zeros(n + 1, dtype=np.int64)
    ip[1:] = np.cumsum([len(r) for r in adj])
    ix = np.concatenate([sorted(r) for r in adj]).astype(np.int64)
    deg = np.diff(ip)
    valid = _all_cuthill_mckee_orders(adj, deg)
    order, _ = OR.rcm_order(ip, ix, reverse=False)
    assert tuple(order.tolist()) in valid


def test_rcm_disconnected_components_and_bandwidth():
    rng = np.random.default_rng(3)
    A = rng.uniform(0, 15, size=(200, 3))
    B = rng.uniform(100, 115, size=(150, 3))
    X = np.concatenate([A, B])
    ip, ix = OG.neighbour_csr(X, 5.0)
    order, undo = OR.rcm_order(ip, ix)
    assert sorted(order.tolist()) == list(range(350))
    assert OR.bandwidth(ip, ix, order) < OR.bandwidth(ip, ix, rng.permutation(350))


# ----------------------------------------------------------------- partition --
def test_batch_head_known_counts():
    # 98,543 atoms with leaves <= 100 gave 1024 batches in the reference's notebook output
    # (Notebook/Application/01_Inching_EigenproblemJDMHD.ipynb cell 4)
    assert len(OP.batch_head(98543, 100)) - 1 == 1024
    for n in (1, 99, 100, 101, 5000, 123457):
        h = OP.batch_head(n, 100)
        sizes = np.diff(h)
        assert h[0] == 0 and h[-1] == n
        assert np.all(sizes <= 100) and np.all(sizes[:-1] >= sizes[1:])
        if n > 100:
            assert np.all(sizes > 25)


def test_islands_cover_all_columns():
    X = small_globule(700, seed=4)
    ip, ix = OG.neighbour_csr(X, 8.0)
    order, _ = OR.rcm_order(ip, ix)
    ip, ix = OG.neighbour_csr(X[order], 8.0)
    isl, half = OP.ungapped_islands(ip, ix, maxleafsize=50, gap_size=20)
    head = OP.batch_head(700, 50)
    assert half == ((ix.size - 700) / 2 + 700) * 9
    for b in range(len(head) - 1):
        cols = np.unique(ix[ip[head[b]]:ip[head[b + 1]]])
        covered = np.zeros(cols.size, bool)
        prev_end = -1
        for (c0, c1) in isl[b]:
            assert c0 < c1 and c0 >= prev_end
            prev_end = c1
            covered |= (cols >= c0) & (cols < c1)
        assert covered.all()
        assert isl[b][0][0] == cols[0] and isl[b][-1][1] == cols[-1] + 1


# Batch 4 of n = 1000 atoms with leaves <= 100 is rows [255, 318) (the halving gives leaf
# sizes 64 x3, 63 x5, 62 x5, 61 x3, sorted descending). Its union of columns has 87
# values; the differences of interest (index iii: left value -> right value):
#   iii = 1  : 11 -> 131 (120 > 100)  index < 5            -> never a gap
#   iii = 6  : 135 -> 252 (117)       island began at 10, 10 + 5 < 255 -> gap [136, 252)
#   iii = 75 : 320 -> 440 (120)       island began at 252, 252 + 5 >= 255 and 321 <= 255 + 63 + 5
#                                      -> the diagonal exception, no gap
#   iii = 78 : 442 -> 580 (138)       443 > 323 -> gap [443, 580)
#   iii = 79 : 580 -> 680 (100)       not > 100 -> no gap
#   iii = 82 : 682 -> 800 (118)       iii == 87 - 5 -> never a gap
# Walking Fuel/Coordinate/T1.py:401-446 by hand gives islands (10,136), (252,443), (580,804).
_KNOWN_UNION = ([10, 11] + list(range(131, 136)) + list(range(252, 321)) + [440, 441, 442, 580, 680, 681, 682]
                + [800, 801, 802, 803])
_KNOWN_ISLANDS = [(10, 136), (252, 443), (580, 804)]


def _known_island_csr():
    """Every row lists only itself, except row 255, which lists the whole designed union."""
    n = 1000
    rows = [[r] for r in range(n)]
    rows[255] = sorted(_KNOWN_UNION)
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    return indptr, np.concatenate(rows).astype(np.int64)


def test_islands_known_answer_hand_derived():
    """Each exception of the reference's gap rule (index < 5, index len - 5, the diagonal
    block, strictly wider than GapSize) is planted once; dropping or loosening any one of
    them changes the batch-4 islands."""
    assert len(_KNOWN_UNION) == 87 and len(set(_KNOWN_UNION)) == 87
    ip, ix = _known_island_csr()
    head = OP.batch_head(1000, 100)
    assert head.tolist() == [0, 64, 128, 192, 255, 318, 381, 444, 507, 569, 631, 693, 755, 817, 878, 939, 1000]
    isl, half = OP.ungapped_islands(ip, ix, maxleafsize=100, gap_size=100)
    assert isl[4] == _KNOWN_ISLANDS
    for b in range(16):
        if b != 4:
            assert isl[b] == [(int(head[b]), int(head[b + 1]))]
    # HalfNnz = ((sum of row lengths - n) / 2 + n) * 9 with sum = 999 + 87
    assert half == 9387.0
    # the same CSR with a gap rule of 99: iii = 79 (580 -> 680, 100 > 99) is now cut as well
    isl99, _ = OP.ungapped_islands(ip, ix, maxleafsize=100, gap_size=99)
    assert isl99[4] == [(10, 136), (252, 443), (580, 581), (680, 804)]


def test_islands_shim_known_answer():
    """The drop-in's island builder (host code over the library's batch partition) on the
    same hand-built CSR, and slice form whatever SliceForm says, as the reference returns."""
    from InchingLiteInteger.Fuel.Coordinate.T1 import ungapped_islands
    ip, ix = _known_island_csr()
    isl, half = ungapped_islands(ip, ix, 1000, maxleafsize=100, gap_size=100)
    assert isl[4] == _KNOWN_ISLANDS and half == 9387.0


# ------------------------------------------------------------------- Hessian --
def test_two_atom_hessian_closed_form():
    X = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 2.0]])
    H = OH.hessian_dense(X, 5.0, plus_i=0.0)
    e = X[1] - X[0]
    E = np.outer(e, e) / (e @ e)
    ref = np.block([[E, -E], [-E, E]])
    assert np.allclose(H, ref, rtol=0, atol=1e-15)


def test_hessian_invariants():
    X = small_globule(400, seed=5)
    A = OH.hessian_dense(X, 8.0, plus_i=1.0)
    H = A - np.eye(A.shape[0])
    assert np.allclose(H, H.T, atol=1e-14)
    Q = OH.rigid_body_basis(X)
    assert np.allclose(Q @ Q.T, np.eye(6), atol=1e-12)
    assert np.max(np.abs(H @ Q.T)) < 1e-11
    ip, ix = OG.neighbour_csr(X, 8.0)
    n_pairs = (ix.size - 400) // 2
    assert np.isclose(np.trace(H), 2 * n_pairs, rtol=1e-12)
    w = np.linalg.eigvalsh(H)
    assert w[0] > -1e-10 and np.sum(np.abs(w) < 1e-9) == 6


def test_periodic_hessian_invariants():
    """Periodic x/y images: symmetric, translations in the null space, exactly three zero
    modes (rotations are no longer free), multi-image entries summed when the box is below
    twice the cutoff; with a box far larger than the structure it equals the open Hessian."""
    rng = np.random.default_rng(8)
    X = rng.uniform(0, 15, (70, 3))
    X[:, 2] *= 0.5
    rp, col, v, m = OH.block_hessian_pbc(X, 8.0, 0.0, (15.0, 15.0))
    n = 70
    D = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            D[3 * i:3 * i + 3, 3 * col[e]:3 * col[e] + 3] = v[e]
    assert np.abs(D - D.T).max() <= 1e-13
    for a in range(3):
        t = np.zeros(3 * n)
        t[a::3] = 1.0
        assert np.abs(D @ t).max() <= 1e-12
    w = np.linalg.eigvalsh(D)
    assert np.sum(np.abs(w) < 1e-9) == 3 and w.min() > -1e-9
    assert any(bin(int(x)).count("1") > 1 for x in m)
    rp2, col2, v2, m2 = OH.block_hessian_pbc(X, 8.0, 1.0, (1000.0, 1000.0))
    rp3, col3, v3 = OH.block_hessian(X, 8.0, 1.0)
    assert np.array_equal(rp2, rp3) and np.array_equal(col2, col3) and np.all(m2 == 1 << 4)
    assert np.max(np.abs(v2 - v3.reshape(v2.shape))) <= 1e-14


def test_lower_csr_and_halfmem_matvec():
    X = small_globule(300, seed=6)
    L, d = OH.lower_csr_and_diag(X, 8.0, plus_i=1.0)
    A = OH.hessian_dense(X, 8.0, plus_i=1.0)
    assert np.allclose(L.toarray(), np.tril(A), atol=0)
    assert np.allclose(d, np.diag(A))
    v = np.random.default_rng(0).standard_normal(900)
    assert np.allclose(OH.halfmem_matvec(L, d, v), A @ v, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- Chebyshev --
def _golden():
    g = np.load(os.path.join(GOLD, "cheb_design.npz"))
    return g, json.loads(bytes(g["meta"]).decode())


def test_chebyshev_design_matches_reference_golden():
    g, meta = _golden()
    for ci, m in enumerate(meta):
        maxd, mind, ker, ext, wanted, bound, ratio = m["case"]
        pol = OC.optimize_on_degree(maxd, mind, ker, ext, tuple(wanted), tuple(bound), ratio)
        assert pol.AdjustedDegree == m["degree"]
        assert np.array_equal(np.asarray(pol.AdjustedCoefficients), g["mu_%d" % ci])
        assert np.array_equal(pol.AdjustedDampingKernels, g["jac_%d" % ci])
        assert float(pol.gam) == m["gam"] and float(pol.cc) == m["cc"] and float(pol.dd) == m["dd"]


def test_damping_kernel_closed_form():
    g, _ = _golden()
    for m in (5, 17, 100):
        k = np.arange(1, m + 1)
        th = np.pi / (m + 2)
        jack = ((1 - k / (m + 2)) * np.sin(th) * np.cos(k * th) + np.cos(th) * np.sin(k * th) / (m + 2)) / np.sin(th)
        # the reference codes a1*sin((k+1)th)/sin th + (1-(k+1)a1) cos(k th); compare at tolerance
        coded = np.sin((k + 1) * th) / ((m + 2) * np.sin(th)) + (1 - (k + 1) / (m + 2)) * np.cos(k * th)
        assert np.allclose(g["damp_Jackson_%d" % m][1:m + 1], coded, rtol=1e-13, atol=1e-15)
        assert g["damp_Jackson_%d" % m][0] == 0.5
        sig = np.sin(k * np.pi / (m + 1)) / (k * np.pi / (m + 1))
        assert np.allclose(g["damp_LanczosSigma_%d" % m][1:m + 1], sig, rtol=1e-14)
        assert np.all(g["damp_None_%d" % m][1:] == 1.0)
        assert jack.shape == coded.shape
        assert np.array_equal(OC.damping_kernel(m, 200, "Jackson"), g["damp_Jackson_%d" % m])


def test_cheb_iv_equals_chebval():
    rng = np.random.default_rng(0)
    mu = rng.standard_normal(30)
    x = np.linspace(-1, 1, 41)
    assert np.allclose(OC.cheb_iv(29, mu, x), np.polynomial.chebyshev.chebval(x, mu), rtol=1e-12, atol=1e-12)


def test_interior_filter_response():
    pol = OC.optimize_on_degree(1000, 10, "Jackson", 1e-9, (5.0, 7.0), (1.0, 40.0), 0.1)
    m = pol.AdjustedDegree
    mu = np.asarray(pol.AdjustedCoefficients)[:m + 1]
    xs = (np.array([5.0, 7.0]) - pol.cc) / pol.dd
    val_edges = np.polynomial.chebyshev.chebval(xs, mu)
    val_centre = np.polynomial.chebyshev.chebval(pol.gam, mu)
    assert np.isclose(val_centre, 1.0, rtol=1e-10)
    assert np.all(val_edges <= 0.1 + 1e-12)


def _small_problem(n_atoms=400, seed=7):
    X = small_globule(n_atoms, seed=seed)
    A = OH.hessian_dense(X, 8.0, plus_i=1.0)
    return X, A


def test_cheb_av_equals_spectral_filter():
    X, A = _small_problem(200, 8)
    w, U = np.linalg.eigh(A)
    pol = OC.optimize_on_degree(500, 5, "Jackson", 1e-10, (1.0, 3.0), (1.0, w[-1] + 1e-12), 0.7)
    v = np.random.default_rng(1).standard_normal(A.shape[0])
    y = OC.cheb_av(lambda x: A @ x, v, pol)
    m = pol.AdjustedDegree
    p = np.polynomial.chebyshev.chebval((w - pol.cc) / pol.dd, np.asarray(pol.AdjustedCoefficients)[:m + 1])
    assert np.allclose(y, U @ (p * (U.T @ v)), rtol=1e-9, atol=1e-9)


def test_kpm_count_is_close_to_exact():
    X, A = _small_problem(300, 9)
    w = np.linalg.eigvalsh(A)
    pol = OC.optimize_on_degree(2000, 5, "Jackson", 1e-10, (1.0, 2.0), (1.0, w[-1] + 1e-12), 0.7)
    est = OC.kpm_estimate(lambda x: A @ x, A.shape[0], pol, 40, np.random.default_rng(0))
    exact = int(np.sum((w >= 1.0 - 1e-9) & (w <= 2.0)))
    assert abs(est - exact) <= max(10, 0.35 * exact)


@pytest.mark.parametrize("designated", [None, 4.0, 1.5])
def test_optimize_on_memory_interval_holds_wanted_count(designated, monkeypatch):
    """The bisection of PolynomialFilters/T2.py:142-257 on a matrix whose spectrum is known
    in closed form (diagonal, eigenvalues 1 + 59 t^2 on a uniform grid of t): the interval it
    settles on must hold the wanted number of eigenvalues to within the margin plus the KPM
    noise of 120 probes (about 1 eigenvalue here). The designated start 4.0 holds 2.3x the
    wanted count (138 eigenvalues) and becomes the right end of the search; 1.5 holds fewer
    than twice the wanted count (55), so the search starts from the spectrum bound: both
    branches of the start rule run."""
    n, wanted, margin = 600, 60, 6
    lam = 1.0 + 59.0 * np.linspace(0.0, 1.0, n) ** 2
    rng = np.random.default_rng(11)
    rights = []
    design = OC.optimize_on_degree

    def spy(*a, **kw):
        rights.append(float(kw["User_WantedInterval"][1]))
        return design(*a, **kw)

    monkeypatch.setattr(OC, "optimize_on_degree", spy)
    pol, cnt, interval = OC.optimize_on_memory(
        lambda x: lam * x, n, rng, User_MaximumDegree=400, User_MinimumDegree=5,
        User_WantedInterval=(1.0, 2.0), User_SpectrumBound=(1.0, 60.0 + 1e-12), User_DesignatedStart=designated,
        User_WantedNumberEigenvalue=wanted, User_AffordableMemoryMargin=margin, User_NumberKpmTrials=120,
        User_ConvergenceRatio=0.7)
    assert abs(cnt - wanted) <= margin                      # the loop's own exit test
    exact = int(np.sum((lam >= interval[0]) & (lam <= interval[1])))
    assert abs(exact - wanted) <= margin + 3, (exact, interval)
    assert interval[0] == 1.0 and interval[1] < (designated if designated == 4.0 else 60.0)
    # the first design is at the designated start; the bisection then starts from [1, 4]
    # or from [1, 60] (start 1.5, and no start)
    first_mid = {None: 30.5, 4.0: 2.5, 1.5: 30.5}[designated]
    expect = [first_mid] if designated is None else [designated, first_mid]
    assert rights[:len(expect)] == pytest.approx(expect, rel=1e-12)
    # the returned filter is the design of the returned interval
    ref = design(400, 5, "Jackson", 1e-10, interval, (1.0, 60.0 + 1e-12), 0.7)
    assert np.array_equal(np.asarray(pol.AdjustedCoefficients), np.asarray(ref.AdjustedCoefficients))


# -------------------------------------------------------------- eigensolvers --
def test_spectrum_bound_below_and_close():
    X, A = _small_problem(300, 10)
    w = np.linalg.eigvalsh(A)
    lo, hi = OL.spectrum_bound(lambda x: A @ x, A.shape[0], np.random.default_rng(0), 200)
    assert hi <= w[-1] * (1 + 1e-10) and hi >= w[-1] * (1 - 1e-6)
    assert lo >= w[0] - 1e-8


def _check_pairs(A, lam, V, w_ref, U_ref, k, tol_val, tol_res):
    assert np.allclose(lam, w_ref[:k], rtol=0, atol=tol_val)
    R = A @ V - V * lam[None, :]
    assert np.max(np.linalg.norm(R, axis=0)) < tol_res
    # the subspace of all returned modes matches (clusters resolved as a whole)
    s = subspace_sines(U_ref[:, :k], V / np.linalg.norm(V, axis=0))
    assert s.max() < 1e-4


def test_trlm_hotelling_matches_dense_eigh():
    X, A = _small_problem(400, 11)
    Q = OH.rigid_body_basis(X)
    w, U = np.linalg.eigh(A)
    k = 12
    lam, V, it = OL.trlm(lambda x: A @ x, A.shape[0], k, np.random.default_rng(0), tol=1e-10, Q=Q, shift=10.0)
    _check_pairs(A, lam, V, w[6:], U[:, 6:], k, 1e-9, 1e-8)


def test_trlm_filtered_matches_dense_eigh():
    X, A = _small_problem(400, 12)
    Q = OH.rigid_body_basis(X)
    w, U = np.linalg.eigh(A)
    k = 10
    lam_max = w[-1] + 1e-12
    pol = OC.optimize_on_degree(2000, 5, "Jackson", 1e-10, (1.0, w[6 + k + 6]), (1.0, lam_max), 0.7)
    lam, V, it = OL.trlm(lambda x: A @ x, A.shape[0], k, np.random.default_rng(0), tol=1e-9, Q=Q, shift=-40.0,
                         pol=pol)
    _check_pairs(A, lam, V, w[6:], U[:, 6:], k, 1e-9, 1e-7)


def test_jdm_matches_dense_eigh():
    X, A = _small_problem(300, 13)
    Q = OH.rigid_body_basis(X)
    w, U = np.linalg.eigh(A)
    k = 8
    lam, V, it = OL.jdm(lambda x: A @ x, A.shape[0], k, np.random.default_rng(0), tol=1e-9, maxiter=2000, Q=Q,
                        shift=40.0, gap=0.0)
    _check_pairs(A, lam, V, w[6:], U[:, 6:], k, 1e-9, 1e-8)


def test_irlm_matches_dense_eigh():
    X, A = _small_problem(400, 14)
    Q = OH.rigid_body_basis(X)
    w, U = np.linalg.eigh(A)
    k = 32
    lam, V, it = OL.irlm(lambda x: A @ x, A.shape[0], k, np.random.default_rng(0), tol=1e-10, Q=Q, shift=10.0,
                         maxiter=3000)
    assert np.allclose(lam, w[6:6 + k], atol=1e-7)


# ---------------------------------------------------------- warm-start network --
def test_oracle_layer_backward_finite_differences():
    """Pins the oracle's hand-derived backward (float64, no rounding) by central differences."""
    X = small_globule(120, seed=72)
    ip, ix = OG.neighbour_csr(X, 8.0)
    rng = np.random.default_rng(1)
    h = rng.standard_normal((120, 4))
    Ws, Wn, b = rng.standard_normal((4, 5)), rng.standard_normal((4, 5)), rng.standard_normal(5)
    dout = rng.standard_normal((120, 5))
    ident = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    f16, ON.f16 = ON.f16, ident
    try:
        out, cache = ON.layer_forward(ip, ix, h, Ws, Wn, b)
        dh, dWs, dWn, db = ON.layer_backward(ip, ix, cache, Ws, Wn, dout)
        L = lambda *a: float(np.sum(ON.layer_forward(ip, ix, *a)[0] * dout))  # noqa: E731
        e = 1e-6
        for (arr, grad) in ((h, dh), (Ws, dWs), (Wn, dWn), (b, db)):
            for idx in [(0,) * arr.ndim, tuple(s - 1 for s in arr.shape)]:
                old = arr[idx]
                arr[idx] = old + e
                lp = L(h, Ws, Wn, b)
                arr[idx] = old - e
                lm = L(h, Ws, Wn, b)
                arr[idx] = old
                assert abs((lp - lm) / (2 * e) - grad[idx]) < 1e-5 * max(1.0, abs(grad[idx]))
    finally:
        ON.f16 = f16
