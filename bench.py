#!/usr/bin/env python
"""Time-to-k-modes benchmark on one MI355X.

Workload (default): the 2.4M-atom assembly -- 60 seeded 40,000-atom globular subunits
(0.058 atoms/A^3, 2.2 A minimum separation, seeds 100..159) rotated by the 60 proper
rotations of the icosahedron onto an orbit of radius 450 A; ANM cutoff 8 A, gamma 1,
PlusI 1; the lowest 1000 non-trivial modes at relative residual
||A v - lam v|| / (lam_max ||v||) <= 1e-8 (float64 check). The subunits are farther apart
than the cutoff, so H has 60 diagonal blocks: each block's rigid-body modes are deflated and
Chebyshev-filtered thick-restart Lanczos runs over all 60 blocks in lockstep.

One step = one full solve from coordinates resident in HBM: neighbour search, RCM ordering
(which also finds the 60 components), cluster layout, spectrum bound, KPM-sized Chebyshev
filter, filtered lockstep TRLM, float64 check. ``value`` is the mean wall time per step
(seconds, lower is better) of the headline operator mode; the plain float64 solver (stored
float64 blocks) and the mixed-precision solve (stored float32 blocks, then float64
refinement) are timed once each in the same run and reported beside it.

``roofline`` describes the kernel with the largest total device time in an instrumented
solve of the headline mode (HIP events on the launch stream around its launches);
``roofline_block_spmv_f32`` the float32 block SpMV (the stored-value operator of the mixed
mode) on the same matrix. ``cpu_baseline`` solves one subunit with the CPU oracle and
ARPACK (the reference's CPU method) and multiplies by the number of subunits.

``--workload globule`` runs the 300,000-atom single-globule line; batch / warm / cpu_plain
the other configurations.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--atoms", type=int, default=300000)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--mode", default="mf64", help="headline operator mode: mf64 | mixed | f64")
    ap.add_argument("--compare", default="f64,mixed,warm,mf64_rows",
                    help="other modes timed once each (warm: the headline operator from the learned start; "
                         "mf64_rows: the headline operator with its recurrence vectors in the matrix's order)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-atoms", type=int, default=20000, help="atoms of the CPU baseline sample")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--warm-atoms", default="300000,1000000", help="globule sizes of --workload warm")
    ap.add_argument("--workload", default="assembly", choices=["assembly", "globule", "batch", "warm", "cpu_plain"],
                    help="assembly: the headline line; globule / batch / warm / cpu_plain: the other "
                         "configurations, one JSON line each")
    ap.add_argument("--modes", type=int, default=1000, help="modes of the assembly workload")
    ap.add_argument("--spmv-only", type=int, default=0, metavar="REPS",
                    help="only build the matrix and launch each operator kernel REPS times (profiling runs)")
    ap.add_argument("--dense-only", type=int, default=0, metavar="REPS",
                    help="only launch the float64 MFMA kernels (Gram, Rayleigh-Ritz combinations) REPS times at "
                         "the solvers' shapes (profiling runs)")
    return ap.parse_args()


def spmv_only(args):
    """Short program for counter collection: the operator kernels of one solve, launched
    REPS times each on the workload's matrix (the assembly, or the globule with
    ``--workload globule``), plus the streaming-copy reference kernel."""
    import torch

    from inching_amd import ops, synth
    from inching_amd.hessian import build_hessian
    if args.workload == "assembly":
        X, _ = synth.assembly(n_sub=60, atoms_per_sub=40000, shell_radius=450.0, seed0=100)
    else:
        X = synth.globule(args.atoms, seed=1)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f32", "f64"))
    n = A.shape[0]
    out = {"workload": args.workload, "nnzb": A.nnzb, "n_atoms": A.n_atoms}
    x32 = torch.randn(n, dtype=torch.float32, device="cuda")
    z32, y32, acc32 = torch.randn_like(x32), torch.empty_like(x32), torch.zeros_like(x32)
    x64 = torch.randn(n, dtype=torch.float64, device="cuda")
    z64, y64, acc64 = torch.randn_like(x64), torch.empty_like(x64), torch.zeros_like(x64)
    from inching_amd.autotune import best_config, bsr_launch
    cfg = {p: best_config(A, p) for p in ("f32", "f64", "mf64")}       # the launch shapes the solver uses
    l32, l64 = bsr_launch(A, cfg["f32"]), bsr_launch(A, cfg["f64"])
    out["kernels"] = {p: kernel_key(p, A) for p in cfg}
    out["bytes_per_launch"] = {kernel_key(p, A): int(kernel_bytes(p, A, "cheb")) for p in cfg}
    if A.clusters() is not None:
        out["cluster_list_entries"] = int(A.clusters()[3].numel())
    cases = {
        "bsr_f32_cheb": lambda: ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, x32, y32, s1=0.1, c0=2.0, z=z32,
                                             c2=1.0, acc=acc32, mu=0.01, lanes=l32[0], col16=l32[1]),
        "bsr_f64_cheb": lambda: ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals64, x64, y64, s1=0.1, c0=2.0, z=z64,
                                             c2=1.0, acc=acc64, mu=0.01, lanes=l64[0], col16=l64[1]),
        "matfree_f64_cheb": lambda: A.matfree_spmv(x64, y64, s1=0.1, c0=2.0, z=z64, c2=1.0, acc=acc64, mu=0.01,
                                                   config=cfg["mf64"]),
    }
    if A.records() is not None:
        (r0, r1, r2), rent = A.records()
        cl = A.clusters()
        acc_r = torch.zeros(3 * cl[0].numel(), dtype=torch.float64, device="cuda")
        ops.rec_pack(cl[0], x64, r0)
        ops.rec_pack(cl[0], z64, r1)
        cases["matfree_rec_f64_cheb"] = lambda: ops.matfree_rec(r0, r1, r2, cl[1], rent, A.plus_i, s1=0.1, c0=2.0,
                                                                c2=1.0, acc=acc_r, mu=0.01)
        out["bytes_per_launch"]["k_matfree_rec<true>"] = int(kernel_bytes("mf64rec", A, "cheb"))
    for name, fn in cases.items():
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.spmv_only):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name + "_avg_ms"] = round(e0.elapsed_time(e1) / args.spmv_only, 5)
    src = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.spmv_only):
        ops.stream_copy(src, dst)
    torch.cuda.synchronize()
    out["stream_copy_bytes_read"] = src.numel() * 8
    from inching_amd.autotune import table
    out["autotune_ms"] = {k: v["ms"] for k, v in table().items()}
    print(json.dumps(out))


def dense_only(args):
    """Short program for counter collection: the MFMA Gram and Rayleigh-Ritz kernels at the
    shapes of the solves -- the 300k-atom globule's restart (basis of 192 rows, 64 Ritz
    vectors, 900k dof) and the 2.4M-atom assembly's lockstep restart (84 rows, 24 vectors per
    component, 60 components of 120k dof) -- REPS launches each, with their mean times."""
    import numpy as np
    import torch

    from inching_amd import ops
    from inching_amd.batch import SegOps, Segments
    dev = torch.device("cuda", 0)
    out = {}

    def timed(name, fn, by, fl):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.dense_only):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.dense_only
        out[name] = {"ms": round(ms, 4), "GBs": round(by / ms / 1e6, 1), "TFLOPs": round(fl / ms / 1e9, 2)}

    m, k, n = 192, 64, 900000
    V = torch.randn((m, n), dtype=torch.float64, device=dev)
    W = torch.randn((m, k), dtype=torch.float64, device=dev)
    Y = torch.empty((k, n), dtype=torch.float64, device=dev)
    ws = ops.GramWorkspace(k, k, dev)
    timed("rr_combine m192 k64 n900k", lambda: ops.rr_combine(V, m, W, Y), (m + k) * n * 8, 2.0 * m * k * n)
    timed("gram k64 x k64 n900k", lambda: ops.gram(Y, Y, ws=ws), 2 * k * n * 8, 2.0 * k * k * n)
    del V, W, Y
    m, k, S, per = 84, 24, 60, 120000
    seg = Segments(np.arange(S + 1, dtype=np.int64) * per, dev)
    so = SegOps(seg, m + 1, dev)
    V = torch.randn((m, S * per), dtype=torch.float64, device=dev)
    W = torch.randn((S, m, k), dtype=torch.float64, device=dev)
    Y = torch.empty((k, S * per), dtype=torch.float64, device=dev)
    timed("seg_combine m84 k24 S60 n7.2M", lambda: so.combine(V, m, W, Y), (m + k) * S * per * 8, 2.0 * m * k * S * per)
    print(json.dumps(out))


def _pct(a, q):
    import numpy as np
    return round(float(np.percentile(a, q)), 4)


def run_batch(args):
    """512 seeded globules, N log-uniform in [500, 20000] (seed 2), lowest 20 modes each, one
    lockstep launch sequence (HIP-graph-captured restarts); solves/s and latency percentiles."""
    import numpy as np
    import torch

    from inching_amd import synth
    from inching_amd.batch import solve_batch
    sizes = synth.batch_sizes(512, seed=2)
    structs = [synth.globule(int(n), seed=1000 + i) for i, n in enumerate(sizes)]
    solve_batch(structs[:16], k=20)                      # warm-up: code paging, allocator, autotune
    torch.cuda.synchronize()
    r = solve_batch(structs, k=20, precision="mf64", tol_rel=1e-8)
    print(json.dumps({"metric": "batched lowest-20-modes solves/s (512 globules, 500-20000 atoms)",
                      "value": round(512 / r.seconds, 3), "unit": "solves/s", "seconds": round(r.seconds, 3),
                      "p50_latency_s": _pct(r.latency, 50), "p99_latency_s": _pct(r.latency, 99),
                      "total_atoms": int(sizes.sum()), "restarts": r.restarts, "converged": r.converged,
                      "max_rel_res": float(np.max(r.rel_res)), "ncv": r.ncv, "operator": "mf64",
                      "lanczos_steps": r.n_spmv}))


def _kernel_stats(timer):
    """Per-kernel totals from HIP event pairs on the launch stream: one pair around each
    Chebyshev filter application (its operator launches run back to back, no other kernel
    between them) and one around every other operator launch."""
    per = {}
    for kind, epi, e0, e1, launches in timer:
        per.setdefault((kind, epi), [0.0, 0])
        per[(kind, epi)][0] += e0.elapsed_time(e1)
        per[(kind, epi)][1] += launches
    return per


def _roofline(A, kind, epi, ms_total, cnt, step_s, workload):
    """Roofline object of the operator kernel (kind, epi) of matrix A: algorithmic bytes per
    launch over its mean launch time, against the HBM peak; for the matrix-free operator
    also its useful float64 rate (18 flops per stored-equivalent 3x3 block) against the
    float64 vector peak."""
    avg_ms = ms_total / cnt
    by = kernel_bytes(kind, A, epi)
    ach = by / (avg_ms * 1e-3) / 1e9
    key = kernel_key(kind, A)
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(key, workload),
           "kernel": key + " (" + KERNEL_DESC[kind] + ")", "epilogue": epi, "avg_launch_ms": round(avg_ms, 5),
           "launches": cnt, "bytes_per_launch": int(by),
           "kernel_share_of_step": round(ms_total / 1e3 / step_s, 3) if step_s else None}
    if kind in ("mf64", "mf64rec"):
        tf = 18.0 * A.nnzb / (avg_ms * 1e-3) / 1e12
        out["useful_fp64"] = {"achieved_tflops": round(tf, 2), "peak_tflops": FP64_VECTOR_PEAK_TFLOPS,
                              "frac": round(tf / FP64_VECTOR_PEAK_TFLOPS, 4),
                              "flops_per_launch": int(18 * A.nnzb),
                              "valu_issue_frac": measured_valu_issue(key, workload)}
    return out


def _block_spmv_f32_roofline(A, reps, workload):
    """The tuned float32 block SpMV with the Chebyshev epilogue on matrix A, ``reps``
    launches bracketed by HIP events on the launch stream."""
    import torch

    from inching_amd import ops
    from inching_amd.autotune import best_config, bsr_launch
    A.ensure("f32")
    n = A.shape[0]
    x = torch.randn(n, dtype=torch.float32, device=A.device)
    z, y, acc = torch.randn_like(x), torch.empty_like(x), torch.zeros_like(x)
    lanes, c16 = bsr_launch(A, best_config(A, "f32"))

    def launch():
        ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, A.vals32, x, y, s1=0.1, c0=2.0, z=z, c2=1.0, acc=acc, mu=0.01,
                     lanes=lanes, col16=c16)
    for _ in range(3):
        launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    return _roofline(A, "f32", "cheb", e0.elapsed_time(e1), reps, None, workload)


def _stream_copy_gbs():
    import torch

    from inching_amd import ops
    src = torch.empty(1 << 28, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        ops.stream_copy(src, dst)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        ops.stream_copy(src, dst)
    e1.record()
    torch.cuda.synchronize()
    return 2 * src.numel() * 8 / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9


def run_assembly(args):
    """The headline line: 60 x 40,000-atom subunits on an icosahedral orbit (radius 450 A),
    lowest ``--modes`` modes at rel_res <= ``--tol``. Headline mode timed over the steps; the
    plain float64 solver (stored float64 blocks) and the mixed-precision solver (stored
    float32 blocks, float64 refinement) once each, all with the same filtered lockstep TRLM."""
    import numpy as np
    import torch

    from inching_amd import synth
    from inching_amd.batch import solve_assembly
    dev = torch.device("cuda", 0)
    X, _ = synth.assembly(n_sub=60, atoms_per_sub=40000, shell_radius=450.0, seed0=100)
    xyz = torch.as_tensor(X, device=dev)
    K = args.modes

    net = None

    def step(mode, timer=None):
        nonlocal net
        v0 = None
        if mode == "warm":                  # matrix-free float64 from the learned start vector
            from inching_amd import warmstart as W
            if net is None:
                net = W.WarmStartNet.load()
            v0 = lambda A_, seg: W.warm_start_vector(net, A_, seg)          # noqa: E731
        if mode == "mf64_rows":             # matrix-free float64, recurrence vectors in the matrix's order
            from inching_amd import operator as OP
            OP.RECORDS = False
            try:
                return solve_assembly(xyz, K, precision="mf64", tol_rel=args.tol, timer=timer, seed=0,
                                      verbose=args.verbose)
            finally:
                OP.RECORDS = True
        return solve_assembly(xyz, K, precision="mf64" if mode == "warm" else mode, tol_rel=args.tol,
                              timer=timer, seed=0, verbose=args.verbose, v0=v0)

    for w in range(args.warmup):           # also fills the autotuner's cache for this matrix
        step(args.mode)
        print("bench: warmup %d done" % w, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.time()
    for s_ in range(args.steps):
        r = step(args.mode)
        print("bench: step %d done (%.3f s)" % (s_, r.seconds), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    sec = (time.time() - t0) / args.steps
    A = r.A

    # instrumented solve of the same workload (graphs off, HIP events around operator launches)
    timer = []
    torch.cuda.synchronize()
    ti = time.time()
    ri = step(args.mode, timer=timer)
    torch.cuda.synchronize()
    t_instr = time.time() - ti
    per = _kernel_stats(timer)
    (kind, epi), (ms_total, cnt) = max(per.items(), key=lambda kv: kv[1][0])
    roof = _roofline(ri.A, kind, epi, ms_total, cnt, t_instr, "assembly")
    kernel_table = {"%s/%s" % kk: {"launches": v[1], "ms_per_step": round(v[0], 3), "avg_ms": round(v[0] / v[1], 5)}
                    for kk, v in per.items()}
    print("bench: instrumented solve done", file=sys.stderr, flush=True)
    del ri, timer
    summary = {"n_atoms": A.n_atoms, "dof": 3 * A.n_atoms, "nnzb_full": A.nnzb, "components": r.components,
               "per_component_modes": r.per_component, "rigid_modes_deflated": 6 * r.components,
               "filter_degree": r.degree, "lanczos_steps": int(r.n_spmv), "restarts": int(r.restarts),
               "operator_launches": int(r.operator_launches), "max_rel_res": float(np.max(r.rel_res)),
               "lam_1_minus_plus_i": float(r.eigval[0] - 1.0), "lam_k_minus_plus_i": float(r.eigval[-1] - 1.0),
               "lam_max_bound": float(r.lam_max), "phases_s": {kk: round(v, 3) for kk, v in r.phases.items()},
               "filter": r.filter_info}
    headline = r
    del A

    compare = {}
    roof_block = None
    for mode in [m for m in args.compare.split(",") if m and m != args.mode]:
        torch.cuda.empty_cache()
        print("bench: compare mode %s start" % mode, file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        ta = time.time()
        rr = step(mode)
        torch.cuda.synchronize()
        compare[mode] = {"seconds": round(time.time() - ta, 3), "max_rel_res": float(np.max(rr.rel_res)),
                         "lanczos_steps": int(rr.n_spmv), "restarts": int(rr.restarts),
                         "operator_launches": int(rr.operator_launches), "filter_degree": rr.degree,
                         "max_abs_dlambda_vs_headline": float(np.max(np.abs(rr.eigval - headline.eigval))),
                         "phases_s": {kk: round(v, 3) for kk, v in rr.phases.items()}}
        if rr.phase1_cycles is not None:
            compare[mode]["float32_phase_cycles"] = int(rr.phase1_cycles)
        print("bench: compare mode %s done in %.2f s" % (mode, compare[mode]["seconds"]), file=sys.stderr, flush=True)
        if getattr(rr.A, "vals32", None) is not None and roof_block is None:
            roof_block = _block_spmv_f32_roofline(rr.A, 20, "assembly")
        del rr
    torch.cuda.empty_cache()
    copy_gbs = _stream_copy_gbs()
    for rf in (roof, roof_block):
        if rf is not None:
            rf["stream_copy_GBs"] = round(copy_gbs, 1)
            rf["frac_of_stream_copy"] = round(rf["achieved"] / copy_gbs, 4)
    if "f64" in compare:
        summary["speedup_vs_plain_f64"] = round(compare["f64"]["seconds"] / sec, 3)
    if "mf64_rows" in compare:
        summary["speedup_records_vs_matrix_order"] = round(compare["mf64_rows"]["seconds"] / sec, 3)
    for mode in ("mixed", "warm"):
        if mode in compare and "f64" in compare:
            summary["speedup_%s_vs_plain_f64" % mode] = round(compare["f64"]["seconds"] / compare[mode]["seconds"], 3)

    cpu = None
    if not args.no_cpu:
        from oracle.cpu_baseline import cpu_baseline_assembly
        print("bench: cpu baseline start", file=sys.stderr, flush=True)
        cpu = cpu_baseline_assembly(synth.globule(40000, seed=100), headline.components,
                                    headline.per_component + 6, args.tol, headline.lam_max,
                                    1.25 * float(np.max(headline.eigval)))
    line = {
        "metric": "time-to-k-modes at resid\u22641e-8 (s) + SpMV HBM GB/s; mixed/learned vs plain FP64",
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"mixed": "f32", "f64": "f64", "mf64": "f64", "f32": "f32"}[args.mode],
        "data": "synthetic (seeded random-sequential-addition globules on an icosahedral orbit, no PDB input)",
        "config": dict({"workload": "2.4M-atom icosahedral assembly (60 x 40,000-atom subunits, orbit radius "
                                    "450 A), lowest %d modes at rel_res <= %g, Chebyshev-filtered lockstep TRLM, "
                                    "operator %s" % (K, args.tol, args.mode)}, **summary),
        "roofline": roof,
        "roofline_block_spmv_f32": roof_block,
        "kernels": kernel_table,
        "compare_modes": compare,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def run_cpu_plain(args):
    """The plain method on the CPU (5,000-atom globule, seed 0, 64 modes, tol 1e-9, FP64 CSR
    Hessian, RCM order, rigid-body Hotelling deflation): the oracle's TRLM on the host beside
    this library's solve of the same matrix on the GPU, both checked against dense eigh."""
    import numpy as np
    import scipy.sparse as sp
    import torch

    from inching_amd import synth
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes
    from oracle import geometry as OG
    from oracle import hessian as OH
    from oracle import lanczos as OL
    from oracle import rcm as OR
    X = synth.globule(5000, seed=0)
    t = time.time()
    ip, ix = OG.neighbour_csr(X, 8.0)
    order, _ = OR.rcm_order(ip, ix)
    Xp = X[order]
    rp, col, v = OH.block_hessian(Xp, 8.0, 1.0)
    H = sp.bsr_matrix((v, col.astype(np.int64), rp), shape=(15000, 15000)).tocsr()
    Qc = OH.rigid_body_basis(Xp)
    t_build = time.time() - t
    t = time.time()
    lam_c, V_c, it_c = OL.trlm(lambda x: H @ x, 15000, 64, np.random.default_rng(0), tol=1e-9, Q=Qc, shift=10.0)
    t_cpu = time.time() - t
    print("bench: cpu trlm done in %.1f s" % t_cpu, file=sys.stderr, flush=True)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f64",))
    Q = rigid_body_basis(A.xyz)
    lowest_modes(A, 64, Q=Q, precision="f64", tol_rel=1e-9)
    torch.cuda.synchronize()
    t = time.time()
    r = lowest_modes(A, 64, Q=Q, precision="f64", tol_rel=1e-9)
    torch.cuda.synchronize()
    t_gpu = time.time() - t
    t = time.time()
    w = np.linalg.eigvalsh(H.toarray())
    t_eigh = time.time() - t
    want = w[6:70]
    print(json.dumps({"metric": "plain method, 5,000-atom globule, 64 modes, tol 1e-9", "unit": "s",
                      "cpu_oracle_trlm_s": round(t_cpu, 3), "cpu_oracle_build_s": round(t_build, 3),
                      "cpu_cores": len(os.sched_getaffinity(0)), "gpu_f64_solve_s": round(t_gpu, 4),
                      "dense_eigh_s": round(t_eigh, 2),
                      "cpu_max_abs_dlambda_vs_eigh": float(np.max(np.abs(lam_c - want))),
                      "gpu_max_abs_dlambda_vs_eigh": float(np.max(np.abs(r.eigval - want))),
                      "gpu_max_rel_res": float(np.max(r.rel_res)), "cpu_restarts": int(it_c)}))


def run_warm(args):
    """Learned warm start vs cold start on unseen 300k- and 1M-atom globules (32 modes):
    thick-restart Lanczos started from a combination of the warm block, and Jacobi-Davidson
    started from the whole warm block; operator launches and time to rel_res 1e-8."""
    import numpy as np
    import torch

    from inching_amd import synth, warmstart as W
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes, lowest_modes_jdm, lowest_modes_jdm_warm, lowest_modes_warm
    net = W.WarmStartNet.load()
    out = {"metric": "learned warm start vs cold start: operator launches and time to rel_res 1e-8", "cases": []}
    sizes = [int(s) for s in args.warm_atoms.split(",")]
    for n_atoms, seed in zip(sizes, (4, 5, 6, 7)):
        X = synth.globule(n_atoms, seed=seed)
        A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
        Q = rigid_body_basis(A.xyz)
        lowest_modes(A, 8, Q=Q, precision="mf64")                 # warm-up at this size
        case = {"n_atoms": n_atoms}
        runs = {}
        for name, fn in (("trlm_cold", lambda: lowest_modes(A, 32, Q=Q, precision="mf64", tol_rel=1e-8)),
                         ("trlm_warm", lambda: lowest_modes_warm(A, 32, net, Q=Q, precision="mf64", tol_rel=1e-8)),
                         ("jdm_cold", lambda: lowest_modes_jdm(A, 32, Q=Q, precision="mf64", tol_rel=1e-8)),
                         ("jdm_warm", lambda: lowest_modes_jdm_warm(A, 32, net, Q=Q, precision="mf64",
                                                                    tol_rel=1e-8))):
            torch.cuda.synchronize()
            t = time.time()
            r = fn()
            torch.cuda.synchronize()
            runs[name] = r
            case[name] = {"seconds": round(time.time() - t, 3), "operator_launches": int(r.total_spmv),
                          "max_rel_res": float(np.max(r.rel_res)), "fallback": bool(getattr(r, "fallback", False))}
            print("bench: warm case %d %s done" % (n_atoms, name), file=sys.stderr, flush=True)
        ref = runs["trlm_cold"].eigval
        for name in ("trlm_warm", "jdm_cold", "jdm_warm"):
            case[name]["max_abs_dlambda_vs_trlm_cold"] = float(np.max(np.abs(runs[name].eigval - ref)))
        case["jdm_launch_ratio_warm_over_cold"] = round(runs["jdm_warm"].total_spmv / runs["jdm_cold"].total_spmv, 3)
        out["cases"].append(case)
        del A, Q, runs
        torch.cuda.empty_cache()
    print(json.dumps(out))


def kernel_bytes(kind, A, epi):
    """Algorithmic bytes of one operator launch of the tuned kernel: stored values + vectors
    for the block SpMV (no index bytes); for the matrix-free operator, whose neighbour indices
    and coordinates are the data it streams, the index layout the tuned kernel reads + the
    coordinates + the vectors."""
    from inching_amd.autotune import best_config
    nnzb, n_atoms = A.nnzb, A.n_atoms
    nd = 3 * n_atoms
    if kind == "mf64rec":
        # record form: lists (slot words) + cluster pointers, every slot's record (coordinates
        # and x, 48 B) read, and per slot 24 B of z read, 24 B of output written, 48 B of the
        # accumulator read and written
        _, ptr, _, ent = A.clusters()
        slots = 4 * (ptr.numel() - 1)
        return 4 * ent.numel() + 8 * ptr.numel() + (144 if epi == "cheb" else 72) * slots
    if kind == "f32":
        s, sv = 4, 4 if epi == "cheb" else 8
        return 9 * s * nnzb + (5 if epi == "cheb" else 2) * nd * sv
    if kind == "f64":
        return 9 * 8 * nnzb + (5 if epi == "cheb" else 2) * nd * 8
    vec = (5 if epi == "cheb" else 2) * nd * 8
    flags = best_config(A, kind) // 1000
    if flags & 4:       # cluster layout: cluster_atoms + cl_ptr + base + list entries
        return sum(t.numel() * t.element_size() for t in A.clusters()) + 24 * n_atoms + vec
    if flags & 1:       # 16-bit column offsets + row bases
        return 2 * nnzb + 8 * (n_atoms + 1) + 4 * n_atoms + 24 * n_atoms + vec
    return 4 * nnzb + 8 * (n_atoms + 1) + 24 * n_atoms + vec


KERNEL_DESC = {"f32": "block SpMV, fp32 values, Chebyshev epilogue",
               "f64": "block SpMV, fp64 values, Chebyshev epilogue",
               "mf64": "matrix-free fp64 ANM product, Chebyshev epilogue",
               "mf64rec": "matrix-free fp64 ANM product on cluster-order records, Chebyshev epilogue"}


def kernel_key(kind, A):
    """Template name of the operator kernel the autotuner picked for this matrix."""
    from inching_amd.autotune import best_config
    if kind == "mf64rec":
        return "k_matfree_rec<true>"
    cfg = best_config(A, kind)
    if kind == "mf64":
        c, f = cfg % 1000, cfg // 1000
        if f & 4 and c > 100:
            return "k_matfree_clp<double,%s>" % ("true" if f & 2 else "false")
        if f & 4:
            return "k_matfree_cl<double,%d,%s>" % (c, "true" if f & 2 else "false")
        return "k_matfree_epi<double,%d,%d,%s,%s>" % (c // 10, c % 10, "unsigned short" if f & 1 else "int",
                                                       "true" if f & 2 else "false")
    t = "float" if kind == "f32" else "double"
    if cfg >= 1200:
        return "k_bsr_spmv<%s,%s,%d,unsigned short,true>" % (t, t, cfg - 1200)
    if cfg >= 200:
        return "k_bsr_spmv<%s,%s,%d,unsigned short,false>" % (t, t, cfg - 200)
    return "k_bsr_spmv<%s,%s,%d,int,false>" % (t, t, cfg)


# committed rocprofv3 counter summaries per workload: (FETCH/WRITE traffic, SQ counters)
PROFILES = {"globule": ("spmv_kernels_globule300k_pmc_traffic.txt", "matfree_kernels_globule300k_pmc_sq.txt"),
            "assembly": ("spmv_kernels_assembly2p4M_pmc_traffic.txt", "matfree_kernels_assembly2p4M_pmc_sq.txt")}


def measured_valu_issue(key, workload):
    """Share of the chip's VALU issue slots the kernel ``key`` fills, from the committed
    rocprofv3 SQ-counter summary (SQ_INSTS_VALU x 4 cycles / (SIMDs x duration)), or None."""
    path = os.path.join(ROOT, "profiles", PROFILES[workload][1])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("k_"):
                    name = line[:line.index(">") + 1].replace(", ", ",")
                    if name == key:
                        return float(line[line.index(">") + 1:].split()[5])
    except (OSError, ValueError, IndexError):
        pass
    return None


def measured_traffic(key, workload):
    """HBM bytes per launch of ``key`` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated on a copy kernel), or None when that kernel was not counted."""
    path = os.path.join(ROOT, "profiles", PROFILES[workload][0])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic "):
                    name, value = line[len("traffic "):].strip().rsplit(None, 1)
                    if name == key:
                        return int(value)
    except OSError:
        pass
    return None


def run_globule(args):
    """The 300,000-atom single-globule workload (64 modes): one line with the headline
    mode's time per step, the other modes timed once each, the dominant kernel's roofline
    and the scaled CPU oracle."""
    import numpy as np
    import torch

    from inching_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from inching_amd.build import build
        build(verbose=False)
    from inching_amd import ops, synth
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes

    dev = torch.device("cuda", 0)
    X = synth.globule(args.atoms, seed=1)
    xyz = torch.as_tensor(X, device=dev)
    torch.cuda.synchronize()

    def step(mode, timer=None):
        from inching_amd import operator as OP
        prec = {"mixed": ("f32",), "f64": ("f64",), "mf64": (), "mf64_rows": ()}[mode]
        A = build_hessian(xyz, 8.0, plus_i=1.0, reorder=True, precision=prec)
        Q = rigid_body_basis(A.xyz)
        OP.RECORDS = mode != "mf64_rows"    # mf64_rows: recurrence vectors in the matrix's order
        try:
            r = lowest_modes(A, args.k, Q=Q, precision="mf64" if mode == "mf64_rows" else mode, tol_rel=args.tol,
                             timer=timer, seed=0)
        finally:
            OP.RECORDS = True
        return A, r

    # warm-up (also JIT-free: the library is prebuilt; this pages in code and the allocator)
    for w in range(args.warmup):
        step(args.mode)
        print("bench: warmup %d done" % w, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    results = []
    t0 = time.time()
    for s_ in range(args.steps):
        results.append(step(args.mode))
        print("bench: step %d done" % s_, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.time()
    sec = (t1 - t0) / args.steps
    A, r = results[-1]
    nnzb, n_atoms = A.nnzb, A.n_atoms

    def kernel_stats(timer, steps):
        """Per-kernel totals from HIP event pairs on the launch stream: one pair around each
        Chebyshev filter application (its operator launches run back to back, no other kernel
        between them) and one around every other operator launch."""
        per = {}
        for kind, epi, e0, e1, launches in timer:
            per.setdefault((kind, epi), [0.0, 0])
            per[(kind, epi)][0] += e0.elapsed_time(e1)
            per[(kind, epi)][1] += launches
        return per

    def roofline(per, step_s, steps):
        if not per:
            return None
        (kind, epi), (ms_total, cnt) = max(per.items(), key=lambda kv: kv[1][0])
        avg_ms = ms_total / cnt
        by = kernel_bytes(kind, A, epi)
        ach = by / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel_key(kind, A), "globule"),
                "valu_issue_frac": measured_valu_issue(kernel_key(kind, A), "globule"),
                "kernel": kernel_key(kind, A) + " (" + KERNEL_DESC[kind] + ")",
                "epilogue": epi, "avg_launch_ms": round(avg_ms, 5), "launches_per_step": cnt // steps,
                "bytes_per_launch": int(by), "kernel_share_of_step": round(ms_total / steps / 1e3 / step_s, 3)}

    # instrumented solve of the same workload: HIP events around each filter application and
    # each other operator launch (the timed steps replay captured HIP graphs, where events
    # cannot be recorded between launches)
    timer = []
    torch.cuda.synchronize()
    ti = time.time()
    step(args.mode, timer=timer)
    torch.cuda.synchronize()
    t_instr = time.time() - ti
    per = kernel_stats(timer, 1)
    roof = roofline(per, t_instr, 1)
    kernel_table = {"%s/%s" % k: {"launches": v[1], "ms_per_step": round(v[0], 3),
                                  "avg_ms": round(v[0] / v[1], 5)} for k, v in per.items()}

    # other modes, one timed solve each (graphs on), plus an instrumented block-SpMV solve
    # (the learned start is the warm workload's subject, not a mode of this line)
    compare = {}
    roof_block = None
    for mode in [m for m in args.compare.split(",") if m and m != args.mode and m != "warm"]:
        torch.cuda.synchronize()
        ta = time.time()
        _, rr = step(mode)
        torch.cuda.synchronize()
        print("bench: compare mode %s done" % mode, file=sys.stderr, flush=True)
        compare[mode] = {"seconds": round(time.time() - ta, 4), "max_rel_res": float(np.max(rr.rel_res)),
                         "total_spmv": int(rr.total_spmv),
                         "max_abs_dlambda_vs_headline": float(np.max(np.abs(rr.eigval - r.eigval)))}
        if mode == "mixed":
            tb = []
            torch.cuda.synchronize()
            tt = time.time()
            step(mode, timer=tb)
            torch.cuda.synchronize()
            pb = {k: v for k, v in kernel_stats(tb, 1).items() if k[0] == "f32"}
            roof_block = roofline(pb, time.time() - tt, 1)

    # achievable bandwidth reference: streaming copy of 2 GiB
    src = torch.empty(1 << 28, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        ops.stream_copy(src, dst)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        ops.stream_copy(src, dst)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 8 / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9
    del src, dst
    for rf in (roof, roof_block):
        if rf is not None:
            rf["stream_copy_GBs"] = round(copy_gbs, 1)
            rf["frac_of_stream_copy"] = round(rf["achieved"] / copy_gbs, 4)

    cpu = None
    if not args.no_cpu:
        from oracle.cpu_baseline import cpu_baseline
        cpu = cpu_baseline(synth.globule(args.cpu_atoms, seed=1), args.atoms, nnzb, args.k, r)

    line = {
        "metric": "time-to-k-modes at resid\u22641e-8 (s) + SpMV HBM GB/s; mixed/learned vs plain FP64",
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"mixed": "f32", "f64": "f64", "mf64": "f64"}[args.mode],
        "data": "synthetic (seeded RSA globule, no PDB input)",
        "config": {"workload": "globule %d atoms, lowest %d modes, TRLM+Chebyshev, mode=%s" % (args.atoms, args.k, args.mode),
                   "n_atoms": n_atoms, "dof": 3 * n_atoms, "nnzb_full": nnzb, "k": args.k,
                   "tol_rel": args.tol, "filter_degree": r.degree, "restarts": r.restarts,
                   "total_spmv": int(r.total_spmv), "n_spmv": r.n_spmv,
                   "phases_s": {k: round(v, 4) for k, v in r.phases.items()},
                   "max_rel_res": float(np.max(r.rel_res)), "lam_min": float(r.eigval[0]),
                   "speedup_vs_plain_f64": (round(compare["f64"]["seconds"] / sec, 3) if "f64" in compare else None),
                   "speedup_records_vs_matrix_order": (round(compare["mf64_rows"]["seconds"] / sec, 3)
                                                       if "mf64_rows" in compare else None)},
        "roofline": roof,
        "roofline_block_spmv_f32": roof_block,
        "kernels": kernel_table,
        "compare_modes": compare,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (the solve runs on a single GPU)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.spmv_only:
        return spmv_only(args)
    if args.dense_only:
        return dense_only(args)
    return {"assembly": run_assembly, "globule": run_globule, "batch": run_batch, "warm": run_warm,
            "cpu_plain": run_cpu_plain}[args.workload](args)


if __name__ == "__main__":
    main()
