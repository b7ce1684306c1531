#!/usr/bin/env python
"""Benchmark: multigrid-PCG throughput on one MI355X.

One step = one full solve of the synthetic ellipsoid-union elasticity problem
(337x294x512 voxels, 6 levels, 56,994,411 DOFs, by default) to relative
residual 1e-6 with the mixed-precision solver: an FP32 symmetric cycle with
Chebyshev sweeps (degree 1 on level 1, 2 on level 2, 3 below) on exact
Galerkin coarse operators inside an FP64 flexible PCG; b resident in HBM,
set-up excluded, the solution left on the device.  The reference algorithm in
FP64 (the baseline of the speed-up), the same cycle in FP64 and two other
FP32 cycles are timed the same way beside it.  The level-1 FP64 operators
(the direction update fused with the matvec, the matvec and the Jacobi step)
are then timed with HIP events on the solver's stream over back-to-back
launches, one instrumented solve gives the per-kernel breakdown, the CPU
baseline solves the 169x147x256 problem with the reference algorithm on the
host cores (oracle/cpu_baseline.cpp), and the 64-problem batch is timed
against one-by-one solves (``c4_batch``).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c3|c4]

--config c4 makes the 64-problem batch the benchmark line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def load_package():
    root = REPO / "gmgs-3d_amd"
    spec = importlib.util.spec_from_file_location(
        "gmgs3d_amd", root / "__init__.py", submodule_search_locations=[str(root)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmgs3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name: str, tx: int, zc: int):
    """HBM bytes per launch of the dominant kernel on (tx, zc) column tiles from
    the committed rocprofv3 --pmc summary (line 'traffic_bytes_per_launch
    tx=TX zc=ZC: N'), or None."""
    f = REPO / "profiles" / f"{name}_traffic.txt"
    if not f.exists():
        return None
    key = f"traffic_bytes_per_launch tx={tx} zc={zc}:"
    for line in f.read_text().splitlines():
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(seed, runs, budget_s=60.0):
    """The reference algorithm's FP64 solve on the host cores: the oracle's
    numpy set-up of the 169x147x256 ellipsoid problem of the same generator
    (timed apart), then oracle/cpu_baseline.cpp (C++/OpenMP, all threads
    OMP_NUM_THREADS allows) solving it to rtol 1e-6, up to ``runs`` times
    within ``budget_s``; the median solve time is the sample."""
    sys.path.insert(0, str(REPO))
    from oracle import cpu_baseline as cb
    from oracle import gmgs_oracle as oracle
    pkg = sys.modules["gmgs3d_amd"]
    if not cb.LIB.exists():
        cb.build()
    problem = pkg.workloads.ellipsoids_c2(seed)
    t0 = time.perf_counter()
    m = oracle.build(problem.vox, problem.bbox, num_levels=problem.num_levels, omega=problem.omega)
    oracle.apply_bc(m, problem.loads, problem.fixed)
    oracle.factor_coarsest(m)
    prep = cb.Prepared(m)
    setup_s = time.perf_counter() - t0
    times, its = [], 0
    t_start = time.perf_counter()
    while len(times) < runs and (not times or time.perf_counter() - t_start + times[-1] < budget_s):
        _, its, conv, rel, sec = prep.solve(m.F, 1e-6, 3000)
        times.append(sec)
        log(f"[bench] cpu baseline solve {len(times)}: {its} its, relres {rel:.2e}, {sec:.2f} s")
    return dict(ndofs=m.levels[0].numDOFs, its=its, seconds=float(np.median(times)), runs=len(times),
                setup_s=setup_s, threads=cb.threads(), cpu=cb.cpu_model(), name=problem.name)


def timed_solves(solver, steps, tag):
    """``steps`` solves from b on the device to convergence (graph-replayed
    iterations, solution left on the device); wall time per solve in ms."""
    times, info = [], None
    for k in range(steps):
        t0 = time.perf_counter()
        _, info, _ = solver.solve(raise_on_nonconvergence=False, copy=False)
        times.append((time.perf_counter() - t0) * 1e3)
        log(f"[bench] {tag} step {k}: its {info.its} relres {info.relres:.3e} "
            f"true {info.true_relres:.3e} {times[-1]:.1f} ms (device {info.solve_seconds * 1e3:.1f})")
    return float(np.mean(times)), info


def batch_figures(args, pkg, steps):
    """The 64-problem batch (uneven ellipsoid-union solids, default level rule):
    ``steps`` gmgs_batch_solve calls after one warm-up, against the same
    problems solved one by one.  Returns the measured figures."""
    t0 = time.perf_counter()
    probs = pkg.workloads.batch_c4(args.seed, 64)
    log(f"[bench] generated 64 problems in {time.perf_counter() - t0:.1f}s")
    flags = (pkg.FLAG_NO_INDEX_TABLES | pkg.FLAG_AUTOTUNE | pkg.FLAG_SYM_VCYCLE
             | pkg.FLAG_NO_JACOBI_STEP | pkg.FLAG_CHEBYSHEV | pkg.FLAG_EXACT_GALERKIN)
    t0 = time.perf_counter()
    solvers = []
    for p in probs:
        s = pkg.Solver(p.vox, p.bbox, num_levels=p.num_levels, jacobi_weight=p.omega, tol=1e-6,
                       max_it=args.max_it, precision=pkg.GMGS_PREC_MIXED, flags=flags,
                       vcycle_sweeps=1, vcycle_sweeps_level2=args.level2_degree,
                       vcycle_sweeps_coarse=args.batch_coarse_degree)
        s.set_bc(p.fixed, p.loads)
        solvers.append(s)
    setup_s = time.perf_counter() - t0
    dofs = [3 * s.num_nodes(1) for s in solvers]
    total = int(sum(dofs))
    log(f"[bench] 64 solvers set up in {setup_s:.1f}s: {total} DOFs "
        f"({min(dofs)}..{max(dofs)} per problem)")
    ref = [s.solve(raise_on_nonconvergence=False)[0] for s in solvers]
    out = pkg.batch_solve(solvers)
    worst = max(float(np.abs(U - R).max() / max(np.abs(R).max(), 1e-300))
                for (U, _), R in zip(out, ref))
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        res = pkg.batch_solve(solvers, copy=False)
        times.append((time.perf_counter() - t0) * 1e3)
        its = [i.its for _, i in res]
        log(f"[bench] batch step {k}: {times[-1]:.1f} ms, its {min(its)}..{max(its)}, "
            f"all converged {all(i.converged for _, i in res)}")
    ms_batch = float(np.mean(times))
    t0 = time.perf_counter()
    for s in solvers:
        s.solve(raise_on_nonconvergence=False, copy=False)
    ms_seq = (time.perf_counter() - t0) * 1e3
    log(f"[bench] sequential: {ms_seq:.1f} ms; batch {ms_batch:.1f} ms")
    for s in solvers:
        s.close()
    return dict(total=total, dofs=dofs, ms_batch=ms_batch, ms_seq=ms_seq, worst=worst, setup_s=setup_s)


def run_batch(args, pkg):
    """--config c4: the batch as the benchmark line."""
    b = batch_figures(args, pkg, args.steps)
    total, dofs, ms_batch, ms_seq, worst, setup_s = (b["total"], b["dofs"], b["ms_batch"], b["ms_seq"],
                                                     b["worst"], b["setup_s"])
    line = {
        "metric": "MDOF/s to rel. residual 1e-6, mixed-prec vs FP64 V-cycle; matvec HBM GB/s",
        "value": round(total / (ms_batch * 1e-3) / 1e6, 3),
        "unit": "MDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(ms_batch, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 PCG",
        "data": "synthetic (64 seeded uneven ellipsoid-union solids)",
        "config": {"workload": f"batch64s{args.seed}", "problems": 64, "dofs": total,
                   "dofs_min": int(min(dofs)), "dofs_max": int(max(dofs)), "rtol": 1e-6,
                   "mixed": ("FP32 symmetric cycle, Chebyshev sweeps (degree 1 on level 1, "
                             f"{args.level2_degree} on level 2, {args.batch_coarse_degree} below), "
                             "exact Galerkin coarse operators, no Jacobi step, FP64 flexible PCG"),
                   "sequential_ms": round(ms_seq, 3),
                   "speedup_vs_sequential": round(ms_seq / ms_batch, 3),
                   "batch_vs_individual_max_rel_dev": worst,
                   "setup_s": round(setup_s, 3)},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c2", "c3", "c4"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--max-it", type=int, default=3000)
    ap.add_argument("--matvec-reps", type=int, default=20)
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the comparison solvers (counter-collection runs)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-problem batch figures of the default line")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--coarse-degree", type=int, default=3,
                    help="Chebyshev degree of the headline cycle on levels 3 and below")
    ap.add_argument("--batch-coarse-degree", type=int, default=4,
                    help="Chebyshev degree of the batch's cycle on levels 3 and below")
    ap.add_argument("--level2-degree", type=int, default=2,
                    help="Chebyshev degree of the headline cycle on level 2")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: the solver runs on one GPU; a multi-GPU slab decomposition is not "
              "implemented. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    pkg = load_package()
    wl = pkg.workloads
    if args.config == "c4":
        run_batch(args, pkg)
        return
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        ms = pkg.copy_probe(nb, reps=5)
        log(f"[bench] copy probe {args.copy_probe_mib} MiB: {ms:.4f} ms "
            f"({2 * nb / (ms * 1e-3) / 1e9:.1f} GB/s read+write)")
    t0 = time.perf_counter()
    prob = wl.ellipsoids_c2(args.seed) if args.config == "c2" else wl.ellipsoids_c3(args.seed)
    log(f"[bench] generated {prob.name} in {time.perf_counter() - t0:.1f}s")
    common = dict(num_levels=prob.num_levels, jacobi_weight=prob.omega, tol=1e-6,
                  max_it=args.max_it)
    base_flags = pkg.FLAG_NO_INDEX_TABLES | pkg.FLAG_AUTOTUNE
    sym_flags = base_flags | pkg.FLAG_SYM_VCYCLE | pkg.FLAG_NO_JACOBI_STEP
    cheb_flags = sym_flags | pkg.FLAG_CHEBYSHEV | pkg.FLAG_EXACT_GALERKIN
    cheb_sweeps = dict(vcycle_sweeps=1, vcycle_sweeps_level2=args.level2_degree,
                       vcycle_sweeps_coarse=args.coarse_degree)

    # headline: FP32 symmetric cycle with Chebyshev sweeps (degree 1 on level 1,
    # 2 on level 2, 3 below) on exact Galerkin coarse operators, inside the FP64
    # flexible PCG
    t0 = time.perf_counter()
    s = pkg.Solver(prob.vox, prob.bbox, precision=pkg.GMGS_PREC_MIXED, flags=cheb_flags,
                   **cheb_sweeps, **common)
    s.set_bc(prob.fixed, prob.loads)
    setup_s = time.perf_counter() - t0
    nn = s.num_nodes(1)
    ndofs = 3 * nn
    log(f"[bench] mixed solver set up in {setup_s:.2f}s: {ndofs} DOFs, {s.num_levels} levels")
    U = None
    for _ in range(max(args.warmup, 1)):
        U, _, _ = s.solve(raise_on_nonconvergence=False)
    ms_per_step, info = timed_solves(s, args.steps, "mixed")
    value = ndofs / (ms_per_step * 1e-3) / 1e6

    def other(tag, precision, flags, sweeps=None):
        o = pkg.Solver(prob.vox, prob.bbox, precision=precision, flags=flags,
                       **(sweeps or dict(vcycle_sweeps=2)), **common)
        o.set_bc(prob.fixed, prob.loads)
        Uo, _, _ = o.solve(raise_on_nonconvergence=False)
        ms, inf = timed_solves(o, args.steps, tag)
        o.close()
        return Uo, ms, inf

    if args.headline_only:
        cmp = None
    else:
        # the reference algorithm in FP64 (kernel/Vcycle.m cycle, Jacobi step,
        # Fletcher-Reeves): the baseline of the speed-up
        U64, ms64, info64 = other("fp64", pkg.GMGS_PREC_FP64, base_flags)
        err = float(np.linalg.norm(U - U64) / np.linalg.norm(U64))
        # the same cycle in FP64 (precision-only comparison) and the reference
        # cycle in FP32
        _, ms64s, info64s = other("fp64-cheb", pkg.GMGS_PREC_FP64, cheb_flags, cheb_sweeps)
        _, msmr, infomr = other("mixed-ref", pkg.GMGS_PREC_MIXED, base_flags)
        # the symmetric V(2,2) cycle with the plain Jacobi weight in every sweep
        _, msmj, infomj = other("mixed-sym-jacobi", pkg.GMGS_PREC_MIXED, sym_flags)
        log(f"[bench] mixed vs fp64 rel L2 {err:.2e}; speed-up {ms64 / ms_per_step:.3f} "
            f"(precision only {ms64s / ms_per_step:.3f})")
        cmp = {"its_fp64": info64.its, "converged_fp64": bool(info64.converged),
               "true_relres_fp64": info64.true_relres,
               "fp64_ms": round(ms64, 3),
               "fp64_mdofs": round(ndofs / (ms64 * 1e-3) / 1e6, 3),
               "speedup_vs_fp64": round(ms64 / ms_per_step, 3),
               "fp64_same_cycle_ms": round(ms64s, 3), "its_fp64_same_cycle": info64s.its,
               "speedup_precision_only": round(ms64s / ms_per_step, 3),
               "mixed_reference_cycle_ms": round(msmr, 3), "its_mixed_reference_cycle": infomr.its,
               "mixed_jacobi_sweeps_ms": round(msmj, 3), "its_mixed_jacobi_sweeps": infomj.its,
               "mixed_vs_fp64_rel_l2": err}

    # the level-1 FP64 operators, back-to-back launches between one event pair:
    # the solver's fused direction update + matvec (its largest FP64 kernel),
    # the plain matvec and the reference algorithm's Jacobi step
    mvp_ms = s.time_operator("matvec64_p", args.matvec_reps)
    mv_ms = s.time_operator("matvec64", args.matvec_reps)
    jac_ms = s.time_operator("jacobi64", args.matvec_reps)
    # algorithmic bytes per solid node, one 32-bit flag word each: the fused
    # kernel alternates between a launch that reads p_old, z (FP32), x and
    # p_2 and writes q, p and x (3 x (8 + 4 + 8 + 8) + 4 read, 3 x 3 x 8
    # written) and one that reads p_old and z and writes q and p (3 x (8 + 4)
    # + 4 read, 3 x 2 x 8 written), timed alternating: per launch the mean of
    # the two, 124 B; q = A p reads p and writes q once; the Jacobi step reads
    # r and 1/diag and writes r_a
    alg_bytes_p = ((3 * (8 + 4 + 8 + 8) + 4 + 3 * 3 * 8) + (3 * (8 + 4) + 4 + 3 * 2 * 8)) * nn // 2
    alg_bytes = 2 * 8 * ndofs + 4 * nn
    jac_bytes = 3 * 8 * ndofs + 4 * nn
    achieved_p = alg_bytes_p / (mvp_ms * 1e-3) / 1e9
    achieved = alg_bytes / (mv_ms * 1e-3) / 1e9
    log(f"[bench] matvec64_p {mvp_ms:.4f} ms ({achieved_p:.1f} GB/s), matvec64 {mv_ms:.4f} ms "
        f"({achieved:.1f} GB/s), jacobi64 {jac_ms:.4f} ms ({jac_bytes / (jac_ms * 1e-3) / 1e9:.1f} GB/s)")
    # the autotuner's column tile of the fused kernel (it picks by time; the
    # HBM traffic differs between tile shapes)
    tcfg = s.query("TILE_CONFIG", 1)
    tx64, zc64 = int(tcfg[3]), int(tcfg[4])
    traffic = read_traffic(f"matvec64p_{prob.name}", tx64, zc64)

    # per-kernel breakdown from one instrumented solve (events around each
    # level-1 launch, graphs off)
    s.kernel_timing(True)
    s.solve(raise_on_nonconvergence=False, copy=False)
    kt = s.kernel_times()
    s.kernel_timing(False)

    cpu = None
    if not args.no_cpu_baseline and cmp is not None:
        log("[bench] cpu baseline: oracle set-up, then the C++/OpenMP FP64 reference-algorithm solve")
        cb = cpu_baseline(args.seed, args.cpu_runs)
        cpu = {"value": round(cb["ndofs"] / cb["seconds"] / 1e6, 4), "unit": "MDOF/s",
               "cores": cb["threads"], "kind": "port",
               "sample": (f"full FP64 solve of {cb['name']} ({cb['ndofs']} DOFs, 5 levels, same generator) "
                          f"to rtol 1e-6 with the reference algorithm (Jacobi step, kernel/Vcycle.m cycle, "
                          f"Fletcher-Reeves), oracle/cpu_baseline.cpp on {cb['threads']} OpenMP threads of "
                          f"'{cb['cpu']}': {cb['its']} iterations, median of {cb['runs']} solve(s) "
                          f"{cb['seconds']:.2f} s; numpy set-up {cb['setup_s']:.1f} s not in the value")}
    batch = None
    if not args.no_batch and not args.headline_only:
        log("[bench] the 64-problem batch")
        b = batch_figures(args, pkg, args.batch_steps)
        batch = {"workload": f"batch64s{args.seed}", "problems": 64, "dofs": b["total"],
                 "dofs_min": int(min(b["dofs"])), "dofs_max": int(max(b["dofs"])),
                 "ms_batch": round(b["ms_batch"], 3), "ms_one_by_one": round(b["ms_seq"], 3),
                 "mdofs_batch": round(b["total"] / (b["ms_batch"] * 1e-3) / 1e6, 3),
                 "speedup_vs_one_by_one": round(b["ms_seq"] / b["ms_batch"], 3),
                 "batch_vs_individual_max_rel_dev": b["worst"], "setup_s": round(b["setup_s"], 3)}
    line = {
        "metric": "MDOF/s to rel. residual 1e-6, mixed-prec vs FP64 V-cycle; matvec HBM GB/s",
        "value": round(value, 3),
        "unit": "MDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 PCG",
        "data": "synthetic (seeded ellipsoid-union voxel solid)",
        "config": {"workload": prob.name, "dofs": ndofs, "levels": s.num_levels,
                   "rtol": 1e-6,
                   "mixed": ("FP32 symmetric cycle, Chebyshev sweeps (degree 1 on level 1, "
                             f"{args.level2_degree} on level 2, {args.coarse_degree} below), exact Galerkin coarse operators, no Jacobi step, "
                             "FP64 flexible PCG"),
                   "fp64": "reference algorithm: kernel/Vcycle.m V-cycle, Jacobi step, FP64 PCG",
                   "its_mixed": info.its, "converged_mixed": bool(info.converged),
                   "true_relres_mixed": info.true_relres,
                   **(cmp or {"comparisons": "skipped (--headline-only)"}),
                   "smoother_lmax": [round(s.smoother(li)[0], 5) for li in range(1, s.num_levels)],
                   "level1_prolongation": ("formed by the residual kernel while staging"
                                           if int(tcfg[0]) == 4 else "kernel of its own"),
                   "setup_s": round(setup_s, 3)},
        "roofline": {"kernel": ("KApplyZm<double, double, 5 | 7, 0, TX> (level-1 direction update and matvec, "
                                "p = z + beta p_old, q = A p, alternating between the launch that also "
                                "applies x += alpha_prev p_2 + alpha p_old (5) and the one without (7); "
                                "per launch the mean of a pair)"), "bound": "hbm",
                     "achieved": round(achieved_p, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_p / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "tile_tx_zc": [tx64, zc64], "alg_bytes_per_launch": alg_bytes_p,
                     "avg_launch_ms": round(mvp_ms, 5), "launches": args.matvec_reps},
        "roofline_matvec64": {"kernel": "KApplyZm<double, double, 0, 0, TX> (level-1 matvec q = A p)",
                              "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                              "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(mv_ms, 5)},
        "roofline_jacobi64": {"kernel": "KApplyZm<double, double, 1, 0, TX> (level-1 Jacobi step)",
                              "achieved": round(jac_bytes / (jac_ms * 1e-3) / 1e9, 1),
                              "frac": round(jac_bytes / (jac_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "alg_bytes_per_launch": jac_bytes, "avg_launch_ms": round(jac_ms, 5)},
        "kernels_ms": {k: round(v[0], 5) for k, v in kt.items()},
        "cpu_baseline": cpu,
        "c4_batch": batch,
    }
    print(json.dumps(line), flush=True)
    s.close()


if __name__ == "__main__":
    main()
