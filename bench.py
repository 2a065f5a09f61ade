"""Benchmark: FETI time-to-solution and dual-operator throughput on one MI355X.

A step is one full solve of the workload from inputs resident in HBM:
local dual operators F_i = (K_i^+)_bb and Dirichlet blocks S_D assembled on
the GPU (FP64 MFMA), coarse (G G^T)^-1, projected PCG to relative dual
residual 1e-8, primal recovery.  The one-time host preprocessing and upload
of the inputs happen before the timed region.

Prints one JSON line (see --help for the workload options).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FETI PCG time-to-solution @ rel residual 1e-8; dual-op applies/s (GB/s %peak)"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(spec, iterations, n_sub, threads, sample):
    """Oracle (CPU restatement of the reference path) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    t0 = time.perf_counter()
    t = O.cpu_sample(spec, sample, threads=threads, preconditioner=3)
    wall = time.perf_counter() - t0
    # per-subdomain costs measured with `threads` threads over `sample` subdomains
    setup = t["setup_per_sub"] * n_sub
    per_iter = (t["kplus_per_sub"] + t["sd_apply_per_sub"]) * n_sub
    total = setup + iterations * per_iter
    return {
        "value": total,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sample} of {n_sub} subdomains timed on the host: fix-point-regularised K_i profile "
                   f"Cholesky + Dirichlet S_D (setup), implicit K_i^+ solve + S_D GEMV (per PCG iteration); "
                   f"scaled to {n_sub} subdomains and the GPU run's {iterations} iterations "
                   f"({wall:.1f} s of host wall time); the hybrid cluster problems, projector and vector "
                   f"updates are not included"),
        "setup_s": setup,
        "per_iteration_s": per_iter,
    }


def oracle_residual(spec, u):
    """Relative residual of the GPU's nodal solution in the oracle's FP64 global
    assembly (no factorisation: one element-by-element K u)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return float(O.global_residual(O.undecomposed(spec), u))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="elasticity_512sub_10e3")
    ap.add_argument("--sc", default="fp64", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--sd", default="fp64", choices=["fp64", "fp32", "bf16"],
                    help="storage of the Dirichlet preconditioner S_D (fp64 or the same as --sc)")
    ap.add_argument("--preconditioner", default="dirichlet")
    ap.add_argument("--method", default=None, choices=["total", "hybrid"],
                    help="FETI method (default: the workload's own, hybrid for the 512- and 1000-subdomain cases)")
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rhs", type=int, default=0,
                    help="also solve this many seeded U(-1,1) nodal load cases on the assembled operators")
    ap.add_argument("--device-assembly", action="store_true",
                    help="assemble the K_i values on the GPU (the host builds structure, loads, kernels, B1/B0)")
    ap.add_argument("--profile-symv", action="store_true",
                    help="only: one setup, --apply-reps F applies and a 1 GiB copy (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs one cluster on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding by clusters is designed, not built).", file=sys.stderr)
        sys.exit(2)

    from espreso_amd import core
    from espreso_amd import problems as PB

    spec = PB.ALL[args.workload]
    if args.device_assembly:
        spec = dict(spec, device_assembly=True)
    method = args.method or spec.get("method", "total")
    log(f"bench: workload {args.workload}, method {method}, F_i storage {args.sc}, loading inputs")
    t0 = time.perf_counter()
    g = core.GridFETI(spec, preconditioner=args.preconditioner, precision=1e-8, sc_precision=args.sc,
                      max_iterations=1000, method=method, sd_precision=args.sd)
    t_build = time.perf_counter() - t0
    log(f"bench: problem built on the host in {t_build:.2f} s ({'K_i values on the device' if args.device_assembly else 'K_i values on the host'}); "
        f"{g.n_sub} subdomains, {g.n_lambda} multipliers")
    t_load = None

    def step():
        g.setup()
        return g.solve()

    if args.profile_symv:
        g.setup()
        sizes = g.sizes()
        ms_F, ms_symv = g.time_apply(0, args.apply_reps)
        ms_copy = core.debug_copy(1 << 30, 3)
        st = g.solve()
        print(json.dumps({"profile": "symv", "workload": args.workload, "sc": args.sc, "ms_symv": ms_symv,
                          "ms_apply": ms_F, "symv_bytes": st["symv_bytes"], "copy_bytes": 1 << 30,
                          "ms_copy": ms_copy, "lambda": sizes[1]}), flush=True)
        g.close()
        return

    for w in range(args.warmup):
        st = step()
        if t_load is None:
            t_load = st["time_load"]
        log(f"bench: warmup {w}: {st['time_total']:.3f} s (assembly {st['time_assembly']:.3f}, "
            f"coarse {st['time_coarse']:.3f}, pcg {st['time_pcg']:.3f} / {st['iterations']} it, "
            f"primal {st['time_primal']:.3f}); load {st['time_load']:.1f} s")
    times, stats = [], []
    t_all = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        st = step()
        times.append(time.perf_counter() - ts)
        stats.append(st)
        log(f"bench: step {k}: {times[-1]:.3f} s wall (assembly {st['time_assembly']:.3f}, "
            f"coarse+rhs {st['time_coarse']:.3f}, pcg {st['time_pcg']:.3f} / {st['iterations']} it, "
            f"primal {st['time_primal']:.3f}), final ||Pr|| {st['final_norm']:.3e} tol {st['tol']:.3e}")
    t_all = time.perf_counter() - t_all
    st = stats[-1]
    tts = t_all / args.steps
    u_gpu = g.nodal_solution()  # before any load case replaces the right-hand side

    rhs_report = None
    if args.rhs > 0:
        loads = PB.nodal_load_cases(spec, args.rhs)
        # together: block re-solve (groups of up to 16 cases share each F_i / S_D pass);
        # one untimed call first allocates the group's vectors (kept by the instance)
        g.resolve_block_nodal(loads[:min(args.rhs, 16)])
        t_b = time.perf_counter()
        _, its_b, _ = g.resolve_block_nodal(loads)
        t_b = time.perf_counter() - t_b
        # one after another (espreso_feti_resolve per case), on the first cases
        n_seq = min(args.rhs, 4)
        its = []
        t_r = time.perf_counter()
        for q in range(n_seq):
            g.set_nodal_rhs(loads[q])
            its.append(g.resolve()["iterations"])
        t_r = time.perf_counter() - t_r
        bs = g.block_stats
        rhs_report = {"cases": args.rhs, "s_per_case_block": round(t_b / args.rhs, 6),
                      "block_phases_s": {"loads": round(bs["time_coarse"], 6), "pcg": round(bs["time_pcg"], 6),
                                         "primal": round(bs["time_primal"], 6)},
                      "iterations_block": [int(v) for v in its_b],
                      "s_per_case_single": round(t_r / n_seq, 6), "single_cases_timed": n_seq,
                      "iterations_single": [int(v) for v in its],
                      "note": "seeded U(-1,1) nodal load cases on the assembled operators; block: "
                              "espreso_grid_resolve_block (nodal split and merge, b, e, interior band solves, "
                              "lockstep PCG, primal; group vectors allocated by an untimed first call); single: "
                              "espreso_grid_set_nodal_rhs + espreso_feti_resolve per case"}
        log(f"bench: {args.rhs} load cases together in {t_b:.3f} s ({t_b / args.rhs * 1e3:.1f} ms per case, "
            f"iterations {int(its_b.min())}..{int(its_b.max())}; loads {bs['time_coarse']:.3f} s, "
            f"pcg {bs['time_pcg']:.3f} s, primal {bs['time_primal']:.3f} s); one after another "
            f"{t_r / n_seq * 1e3:.1f} ms per case")

    # dual operator applies on resident vectors, SYMV kernel timed with HIP events
    ms_F, ms_symv = g.time_apply(0, args.apply_reps)
    ms_M, ms_symv_sd = g.time_apply(1, args.apply_reps)
    gbs = st["symv_bytes"] / (ms_symv * 1e-3) / 1e9
    gbs_sd = st["symv_bytes_sd"] / (ms_symv_sd * 1e-3) / 1e9
    log(f"bench: F apply {ms_F:.3f} ms (SYMV {ms_symv:.3f} ms, {gbs:.0f} GB/s); "
        f"S_D apply {ms_M:.3f} ms (SYMV {ms_symv_sd:.3f} ms, {gbs_sd:.0f} GB/s)")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "symv_traffic.json")
    if os.path.exists(tfile):
        try:
            tr = json.load(open(tfile))
            if tr.get("workload") == args.workload and tr.get("sc") == args.sc:
                traffic = tr.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    asm_tflops = st["assembly_flops"] / st["time_assembly"] / 1e12
    # what the FP64 MFMA instruction of the tile routine sustains on this GPU with
    # register operands only (no memory, no LDS): the practical ceiling
    mfma_loop = core.debug_mfma_rate(0, 8)
    result = {
        "metric": METRIC,
        "value": round(tts, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tts * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.sc == "fp64" and args.sd == "fp64" else f"f64 (F_i {args.sc}, S_D {args.sd})",
        "data": "synthetic (seeded structured HEXA8 box)",
        "config": {
            "workload": args.workload,
            "method": "HYBRID_FETI" if method == "hybrid" else "TOTAL_FETI",
            "preconditioner": args.preconditioner,
            "subdomains": g.n_sub,
            "elements_per_subdomain": (list(spec["elements"]) if spec.get("slab_elements") is None
                                       else "uneven: " + " / ".join(f"{min(v)}..{max(v)}" for v in spec["slab_elements"])),
            "lambda": g.n_lambda,
            "coarse": st["n_coarse"],
            "sc_storage": args.sc,
            "sd_storage": args.sd,
            "precision": 1e-8,
            "symv_item_elements": {k: v for k, v in zip(("fp64", "fp32", "bf16"), g.symv_items()) if v},
        },
        "iterations": st["iterations"],
        "phases_s": {k: round(st[k], 6) for k in ("time_assembly", "time_coarse", "time_pcg", "time_primal")},
        "inputs_s": {"host_build": round(t_build, 4), "first_load": None if t_load is None else round(t_load, 4),
                     "K_values": "device" if args.device_assembly else "host",
                     "note": "building and uploading the assembled inputs; outside the timed time-to-solution"},
        "dual_op_applies_per_s": round(1e3 / ms_F, 2),
        "dual_op_GBps": round(gbs, 1),
        "dirichlet_applies_per_s": round(1e3 / ms_M, 2),
        "assembly_TFLOPs": round(asm_tflops, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_symv (F_i packed SYMV with fused lambda gather)",
            "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": st["symv_bytes"],
            "avg_ms": round(ms_symv, 4),
        },
        "roofline_assembly": {
            "bound": "mfma",
            "achieved": round(asm_tflops, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(asm_tflops / FP64_PEAK_TFLOPS, 4),
            "flops": st["assembly_flops"],
            "seconds": round(st["time_assembly"], 4),
            "mfma_loop_TFLOPs": round(mfma_loop, 2),
            "frac_of_mfma_loop": round(asm_tflops / mfma_loop, 4),
            "note": "achieved = executed assembly flops / assembly time; mfma_loop = v_mfma_f64_16x16x4f64 "
                    "issue rate with register operands only, measured in this run (espreso_debug_mfma_rate)",
        },
    }
    if rhs_report:
        result["load_cases"] = rhs_report
    g.close()
    accuracy = {"dual_rel_residual": st["final_norm"] / st["tol"] * 1e-8,
                "dual_note": "||P r|| / min(||P r0||, ||b||) at the stop (stop test < 1e-8)"}
    if args.sc != "fp64" or args.sd != "fp64":
        # the same workload on the FP64 path (whose parity with the oracle the tests establish)
        g64 = core.GridFETI(spec, preconditioner=args.preconditioner, precision=1e-8, max_iterations=1000,
                            method=method)
        g64.setup()
        g64.solve()
        u64 = g64.nodal_solution()
        g64.close()
        accuracy["primal_rel_diff_vs_fp64_gpu"] = float(np.abs(u_gpu - u64).max() / np.abs(u64).max())
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} host threads, sample {args.cpu_sample} subdomains")
        result["cpu_baseline"] = cpu_baseline(spec, st["iterations"], st["n_subdomains"], threads, args.cpu_sample)
        accuracy["primal_rel_residual_oracle"] = oracle_residual(spec, u_gpu)
        accuracy["primal_note"] = ("||K u - f|| / ||f|| over the free DOFs of the undecomposed problem, K and f "
                                   "assembled element by element by the oracle in FP64")
    result["accuracy"] = accuracy
    log("bench: accuracy " + json.dumps(accuracy))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
