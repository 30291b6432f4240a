"""CPU oracle for the eigenproblem construction + diagonalisation path.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or as the timed CPU baseline). The product path (``inching_amd`` and the
``InchingLiteInteger`` shim) never imports it and has no CPU fallback.

It is a fresh numpy/scipy (and plain C, see ``oracle/c``) restatement of the
reference algorithms; every function cites the reference file:line it follows
(paths relative to the reference repository root).

Modules and how each function is pinned (what makes a plausible mistake fail):

geometry.neighbour_csr        brute force all-pairs on small inputs; cKDTree;
                              symmetry; self included.
rcm.rcm_order                 scipy.sparse.csgraph.reverse_cuthill_mckee on the
                              same CSR (descending column order) with the same
                              seed order; brute-force enumeration of all valid
                              Cuthill-McKee orders on tiny graphs; C twin
                              (oracle/c/rcm_oracle.c) agrees bit-for-bit.
rcm.rcm_order_levels          level-synchronous form (the device algorithm):
                              equal to rcm_order on random geometric and
                              non-geometric graphs, both seed orders, both
                              directions.
partition.batch_head          closed form of recursive halving (leaf sizes in
                              (L/2, L], sum = N) on many N.
partition.ungapped_islands    hand-derived islands of a planted CSR whose
                              early-index, fifth-last, diagonal and width
                              exceptions of the gap rule each change the
                              answer; brute-force coverage (every neighbour
                              column of a batch lies in an island).
hessian.*                     closed form -d d^T/|d|^2; invariants (symmetry,
                              translations and infinitesimal rotations in the
                              null space of H, tr H = 2 #contacts, PSD);
                              dense <-> CSR agreement.
hessian.rigid_body_basis      orthonormality Q Q^T = I_6 and H Q^T = 0.
hessian.block_hessian_pbc     periodic x/y images by brute force over all pairs
                              and nine shifts: symmetry, translations in the
                              null space, exactly three zero modes, summed
                              multi-image entries, equal to block_hessian
                              when the box is far larger than the structure.
chebyshev.damping_kernel      closed form (Jackson / Lanczos sigma).
chebyshev.optimize_on_degree  golden vectors written by importing the
                              reference's numpy-only PolynomialFilters/T0.py
                              (tests/golden/make_cheb_golden.py); filter
                              response p(x) = sum mu_k T_k(x) vs numpy chebval.
chebyshev.cheb_av             U p(L) U^T x by dense eigh on small H.
chebyshev.kpm_estimate        statistical only (mean count vs dense eigh count).
chebyshev.optimize_on_memory  closed-form spectrum: the chosen interval holds the
                              wanted count within the margin (both
                              designated-start branches).
lanczos.spectrum_bound        lambda_max estimate <= true lambda_max and close.
analysis.sqfluct              with every non-rigid eigenpair of a small H,
                              sum_m RatioSqFPM^2 = lambda_min tr(H^+)_aa.
analysis.cross_correlation    self-pair maps undone and summed over every
                              eigenpair = 3x3 block traces of H^2; hand-built
                              single-atom modes fix the orientation and sign.
lanczos.trlm / jdm / irlm     final eigenpairs vs dense numpy.linalg.eigh
                              (eigenvalues, subspace angles, residuals).
                              Iterates and iteration counts are parity
                              UNPINNED (they depend on CuPy's RNG stream and
                              cuBLAS reduction order in the reference).
"""
