"""Benchmark: FETI time-to-solution and dual-operator throughput on one MI355X.

A step is one cold solve of the workload: from the per-subdomain inputs handed
to the library (espreso_feti_set_subdomain: K_i, f_i, R_i, B1 rows, c) to the
primal solution on the device.  It covers the preprocessing and upload of
the inputs, the local dual operators F_i = (K_i^+)_bb and Dirichlet blocks S_D
assembled on the GPU (FP64 MFMA), the coarse problem, projected PCG to
relative dual residual 1e-8 and primal recovery.  The structured-grid
builder that produces the inputs (the reference's assembler) runs in the
same call but is measured separately and left out of the step time.

Default workload: 1000 subdomains of 12^3 HEXA8 elasticity elements, TFETI,
Dirichlet preconditioner, F_i and S_D in FP64.  Beside it the line reports
the warm re-solve after new element moduli (espreso_grid_update_material +
set-up + solve on the resident instance), the mixed-precision path (F_i FP32,
S_D BF16, FP64 iterative refinement) with its speed-up over FP64, the primal
error of each path against the oracle's full-size solution (golden fixture),
the SYMV and assembly rooflines and the CPU restatement timed on the host.

Prints one JSON line (see --help).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FETI PCG time-to-solution @ rel residual 1e-8; dual-op applies/s (GB/s %peak)"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
HOST_THREADS_PER_GPU = 16  # the host share of one GPU on the benchmark machines


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(spec, iterations, n_sub, threads, sample):
    """The oracle's explicit-S_i restatement (dense Schur complement and its
    generalised inverse per subdomain, dense F_i and S_D applies) on a sample
    of subdomains, plus the full-size coarse set-up and per-iteration vector
    work, on the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    t0 = time.perf_counter()
    t = O.cpu_baseline(spec, sample, threads=threads)
    wall = time.perf_counter() - t0
    setup = t["setup_per_sub"] * n_sub + t["coarse_setup"]
    per_iter = (t["f_apply_per_sub"] + t["sd_apply_per_sub"]) * n_sub + t["iteration_rest"]
    total = setup + iterations * per_iter
    measured_full = t["coarse_setup"] + iterations * t["iteration_rest"]
    return {
        "value": round(total, 3),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"explicit S_i (dense Schur complement of every subdomain and its generalised inverse F_i, S_D from "
                   f"the same step) timed on {t['sample']} of {n_sub} subdomains with {threads} threads, scaled "
                   f"to {n_sub}; dense F_i and S_D applies per PCG iteration timed on the same sample and scaled; "
                   f"the coarse G G^T set-up and one iteration's projections, B gathers/scatters, dots and updates "
                   f"timed at full size; {iterations} iterations (the GPU run's count, same algorithm); "
                   f"{wall:.1f} s of host wall time"),
        "setup_s": round(setup, 3),
        "per_iteration_s": round(per_iter, 5),
        "extrapolated_fraction": round(1.0 - measured_full / total, 4),
        "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model(),
                 "threads_note": f"{threads} threads = the host share of one GPU on this machine"},
    }


def oracle_residual(spec, u):
    """Relative residual of the nodal solution in the oracle's FP64 global
    assembly (element by element, no factorisation)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return float(O.global_residual(O.undecomposed(spec), u))


def fixture_error(workload, u):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import fullsize
    fx = fullsize.load(workload)
    if fx is None:
        return None
    d = fullsize.compare(fx, u)
    d["reference"] = (f"oracle TFETI solve to relative dual residual {float(fx['precision']):g} "
                      f"(tests/golden/fullsize_{workload}.npz: {len(fx['sample_nodes'])} sampled nodes, 2-norm and "
                      f"a seeded checksum over all DOFs)")
    return d


def phases(st):
    return {k: round(st[k], 6) for k in ("time_load", "time_assembly", "time_coarse", "time_pcg", "time_primal")}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="elasticity_1000sub_12e3")
    ap.add_argument("--method", default=None, choices=["total", "hybrid"],
                    help="FETI method (default: the workload's own)")
    ap.add_argument("--preconditioner", default="dirichlet")
    ap.add_argument("--sc", default="fp64", choices=["fp64", "fp32", "bf16"], help="storage of F_i (value path)")
    ap.add_argument("--sd", default="fp64", choices=["fp64", "fp32", "bf16"], help="storage of S_D (value path)")
    ap.add_argument("--mixed", default="fp32,bf16",
                    help="F_i,S_D storage of the mixed-precision path reported beside (empty: skip)")
    ap.add_argument("--mixed-steps", type=int, default=2)
    ap.add_argument("--warm-steps", type=int, default=2)
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=HOST_THREADS_PER_GPU)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rhs", type=int, default=0,
                    help="also solve this many seeded U(-1,1) nodal load cases on the assembled operators")
    ap.add_argument("--device-assembly", action="store_true",
                    help="K_i values computed on the GPU (the host builds structure, loads, kernels, B1/B0)")
    ap.add_argument("--profile-symv", action="store_true",
                    help="only: one set-up, --apply-reps F applies and a 1 GiB copy (profiler counter runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs one cluster of subdomains on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding by clusters over several GPUs is designed, not built).", file=sys.stderr)
        sys.exit(2)

    from espreso_amd import core
    from espreso_amd import problems as PB

    spec = PB.ALL[args.workload]
    if args.device_assembly:
        spec = dict(spec, device_assembly=True)
    method = args.method or spec.get("method", "total")
    opts = dict(preconditioner=args.preconditioner, precision=1e-8, max_iterations=1000, method=method)

    def cold(sc, sd):
        """One cold solve; returns (instance, stats, step seconds without the host grid builder)."""
        t0 = time.perf_counter()
        g = core.GridFETI(spec, sc_precision=sc, sd_precision=sd, **opts)
        g.setup()
        st = g.solve()
        wall = time.perf_counter() - t0
        return g, st, wall - st["time_build"]

    if args.profile_symv:
        g = core.GridFETI(spec, sc_precision=args.sc, sd_precision=args.sd, **opts)
        g.setup()
        ms_F, ms_symv = g.time_apply(0, args.apply_reps)
        ms_copy = core.debug_copy(1 << 30, 3)
        st = g.solve()
        print(json.dumps({"profile": "symv", "workload": args.workload, "sc": args.sc, "ms_symv": ms_symv,
                          "ms_apply": ms_F, "symv_bytes": st["symv_bytes"], "copy_bytes": 1 << 30,
                          "ms_copy": ms_copy, "lambda": g.sizes()[1]}), flush=True)
        g.close()
        return

    log(f"bench: workload {args.workload}, method {method}, F_i {args.sc}, S_D {args.sd}; cold steps")
    first = None
    for w in range(args.warmup):
        g, st, t = cold(args.sc, args.sd)
        if first is None:
            first = t
        log(f"bench: warmup {w}: {t:.3f} s (build {st['time_build']:.2f} s excluded; load {st['time_load']:.3f}, "
            f"assembly {st['time_assembly']:.3f}, coarse {st['time_coarse']:.3f}, pcg {st['time_pcg']:.3f} / "
            f"{st['iterations']} it, primal {st['time_primal']:.3f})")
        g.close()
    times, stats = [], []
    g = None
    for k in range(args.steps):
        if g is not None:
            g.close()
        g, st, t = cold(args.sc, args.sd)
        times.append(t)
        stats.append(st)
        log(f"bench: step {k}: {t:.3f} s (build {st['time_build']:.2f} s excluded; load {st['time_load']:.3f}, "
            f"assembly {st['time_assembly']:.3f}, coarse {st['time_coarse']:.3f}, pcg {st['time_pcg']:.3f} / "
            f"{st['iterations']} it, primal {st['time_primal']:.3f}), ||Pr|| {st['final_norm']:.3e} tol "
            f"{st['tol']:.3e}")
    st = stats[-1]
    tts = float(np.mean(times))
    u_gpu = g.nodal_solution()

    # warm: new element moduli on the loaded instance (K recomputed on the
    # device), then set-up and solve on resident inputs
    warm = None
    if args.warm_steps > 0 and spec.get("physics", 0) == 0:
        E = np.full(core.n_elements(spec), spec.get("young", PB.E_STEEL))
        wt = []
        for k in range(args.warm_steps):
            t0 = time.perf_counter()
            g.update_material(E)
            g.setup()
            sw = g.solve()
            wt.append(time.perf_counter() - t0)
        warm = {"value_s": round(float(np.mean(wt)), 6), "steps": args.warm_steps, "iterations": sw["iterations"],
                "phases_s": {k: round(sw[k], 6) for k in ("time_assembly", "time_coarse", "time_pcg", "time_primal")},
                "note": "espreso_grid_update_material (K_i recomputed on the GPU from new element moduli) + "
                        "espreso_feti_setup + espreso_feti_solve on the loaded instance: no host preprocessing "
                        "or upload of K"}
        log(f"bench: warm re-solve {warm['value_s']:.3f} s")

    rhs_report = None
    if args.rhs > 0:
        loads = PB.nodal_load_cases(spec, args.rhs)
        g.resolve_block_nodal(loads[:min(args.rhs, 16)])  # allocates the group's vectors (kept by the instance)
        t_b = time.perf_counter()
        _, its_b, _ = g.resolve_block_nodal(loads)
        t_b = time.perf_counter() - t_b
        bs = g.block_stats
        rhs_report = {"cases": args.rhs, "s_per_case_block": round(t_b / args.rhs, 6),
                      "iterations_block": [int(v) for v in its_b],
                      "block_phases_s": {"loads": round(bs["time_coarse"], 6), "pcg": round(bs["time_pcg"], 6),
                                         "primal": round(bs["time_primal"], 6)}}
        log(f"bench: {args.rhs} load cases together: {t_b / args.rhs * 1e3:.1f} ms per case")

    # dual-operator applies on resident vectors; the SYMV kernel timed with HIP
    # events recorded on the stream it runs on
    ms_F, ms_symv = g.time_apply(0, args.apply_reps)
    ms_M, ms_symv_sd = g.time_apply(1, args.apply_reps)
    gbs = st["symv_bytes"] / (ms_symv * 1e-3) / 1e9
    gbs_sd = st["symv_bytes_sd"] / (ms_symv_sd * 1e-3) / 1e9
    log(f"bench: F apply {ms_F:.3f} ms (SYMV {ms_symv:.3f} ms, {gbs:.0f} GB/s); "
        f"S_D apply {ms_M:.3f} ms (SYMV {ms_symv_sd:.3f} ms, {gbs_sd:.0f} GB/s)")
    n_sub, n_lambda = g.n_sub, g.n_lambda
    symv_items = {k: v for k, v in zip(("fp64", "fp32", "bf16"), g.symv_items()) if v}
    g.close()

    # HBM bytes per F SYMV launch from the FETCH_SIZE / WRITE_SIZE counter runs
    # of an earlier profile (counters cannot be read inside this run)
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "symv_traffic.json")
    if os.path.exists(tfile):
        try:
            for tr in json.load(open(tfile)).get("entries", []):
                if tr.get("workload") == args.workload and tr.get("sc") == args.sc:
                    traffic = tr.get("bytes_per_launch")
                    traffic_src = "rocprofv3 FETCH_SIZE + WRITE_SIZE runs, " + tr.get("source", "")
        except (OSError, ValueError, AttributeError):
            traffic = None

    asm_tflops = st["assembly_flops"] / st["time_assembly"] / 1e12
    # the FP64 MFMA ceiling measured in this run: the better of the 16x16x4
    # loop and the 4x4x4 (4-block) loop on an 8 x 8 operand grid, which keeps
    # the matrix pipe closest to its peak of 32 flops per clock and SIMD
    mfma_loop_16 = core.debug_mfma_rate(0, 8)
    mfma_loop_4 = core.debug_mfma_rate(2, 8)
    mfma_loop = max(mfma_loop_16, mfma_loop_4)

    accuracy = {"dual_rel_residual": st["final_norm"] / st["tol"] * 1e-8,
                "dual_note": "||P r|| / min(||P r0||, ||b||) at the stop (stop test < 1e-8)",
                "kkt_primal": st["kkt_primal"], "kkt_dual": st["kkt_dual"],
                "primal_error_vs_oracle": fixture_error(args.workload, u_gpu)}

    # mixed precision beside the FP64 path (cold, the same workload and method)
    mixed = None
    if args.mixed and (args.sc, args.sd) == ("fp64", "fp64"):
        msc, msd = args.mixed.split(",")
        mt, mst = [], None
        for k in range(args.mixed_steps):
            gm, mst, t = cold(msc, msd)
            mt.append(t)
            if k < args.mixed_steps - 1:
                gm.close()
        um = gm.nodal_solution()
        gm.close()
        mtime = float(np.mean(mt))
        mixed = {"sc_storage": msc, "sd_storage": msd, "mixed_time_s": round(mtime, 6),
                 "speedup_vs_fp64": round(tts / mtime, 4), "steps": args.mixed_steps,
                 "iterations": mst["iterations"], "refinement_steps": mst["refinement_steps"],
                 "phases_s": phases(mst),
                 "accuracy": {"dual_rel_residual": mst["final_norm"] / mst["tol"] * 1e-8,
                              "primal_error_vs_oracle": fixture_error(args.workload, um),
                              "primal_rel_diff_vs_fp64_gpu": float(np.abs(um - u_gpu).max() / np.abs(u_gpu).max())},
                 "note": "F_i and S_D assembled in FP64, stored and applied in the given precisions with FP32 "
                         "accumulation; outer PCG, coarse problem, R_i and refinement residual in FP64"}
        log(f"bench: mixed {msc}/{msd}: {mtime:.3f} s, speed-up {tts / mtime:.3f} over FP64 "
            f"(pcg {mst['time_pcg']:.3f} s / {mst['iterations']} it)")

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
        "dtype": "f64" if (args.sc, args.sd) == ("fp64", "fp64") else f"f64 (F_i {args.sc}, S_D {args.sd})",
        "data": "synthetic (seeded structured HEXA8 box)",
        "config": {
            "workload": args.workload,
            "method": "HYBRID_FETI" if method == "hybrid" else "TOTAL_FETI",
            "preconditioner": args.preconditioner,
            "subdomains": n_sub,
            "elements_per_subdomain": (list(spec["elements"]) if spec.get("slab_elements") is None
                                       else "uneven: " + " / ".join(f"{min(v)}..{max(v)}"
                                                                    for v in spec["slab_elements"])),
            "lambda": n_lambda,
            "coarse": st["n_coarse"],
            "sc_storage": args.sc,
            "sd_storage": args.sd,
            "precision": 1e-8,
            "symv_item_elements": symv_items,
        },
        "timed": ("cold: a new instance from the espreso_feti_set_subdomain data to the primal solution on the "
                  "device (handoff of the inputs, their preprocessing and upload, assembly, coarse problem, PCG, "
                  "primal recovery); the structured-grid builder that produces the inputs is measured apart "
                  "(inputs_s.host_build). The process keeps the assembly workspace once allocated and lends it "
                  "to later instances; the first solve of the process also pays that allocation and the loading "
                  "of the kernels (first_solve_in_process_s)"),
        "first_solve_in_process_s": None if first is None else round(first, 6),
        "iterations": st["iterations"],
        "phases_s": phases(st),
        "inputs_s": {"host_build": round(st["time_build"], 4),
                     "K_values": "device" if args.device_assembly else "host"},
        "warm": warm,
        "mixed": mixed,
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
            "traffic_source": traffic_src,
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
            "mfma_loop_16x16x4_TFLOPs": round(mfma_loop_16, 2),
            "mfma_loop_4x4x4_TFLOPs": round(mfma_loop_4, 2),
            "frac_of_mfma_loop": round(asm_tflops / mfma_loop, 4),
            "note": "achieved = executed assembly flops / assembly time; mfma_loop = the higher FP64 MFMA issue rate "
                    "with register operands only of v_mfma_f64_16x16x4f64 (8 chains) and v_mfma_f64_4x4x4f64 "
                    "(8 x 8 operand grid), measured in this run (espreso_debug_mfma_rate)",
        },
    }
    if rhs_report:
        result["load_cases"] = rhs_report
    if not args.no_cpu:
        threads = min(HOST_THREADS_PER_GPU, os.cpu_count() or 1)
        log(f"bench: CPU baseline on {threads} host threads, sample {args.cpu_sample} subdomains")
        result["cpu_baseline"] = cpu_baseline(spec, st["iterations"], n_sub, threads, args.cpu_sample)
        accuracy["primal_rel_residual_oracle"] = oracle_residual(spec, u_gpu)
        accuracy["primal_residual_note"] = ("||K u - f|| / ||f|| over the free DOFs of the undecomposed problem, "
                                            "K and f assembled element by element by the oracle in FP64")
    result["accuracy"] = accuracy
    log("bench: accuracy " + json.dumps(accuracy))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
