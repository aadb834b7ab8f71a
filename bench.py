#!/usr/bin/env python
"""RBF-PUM setup+apply throughput on one MI355X.

A step is one full transfer, setup (patch centres and radii, membership,
weights) plus apply (per-patch solve and blend) of every field, on inputs
already resident in HBM.

Default (--config c3): the uneven-cloud workload the headline metric is
quoted on: 4,096,000 sources from a 12-component Gaussian mixture in
[0, 1]^3 (seed 4) -> 4,096,000 uniform targets (seed 5), density-aware
patches (leaf 40 sources, 20 to 600 per patch, overlap 0.15, support = patch
radius), 8 fields (franke_3d + 7 seeded trig fields), Wendland C2. The timed
step is the batched transfer: one plan, all 8 fields per solve, each size
class of patches in the variant and precision the size-bin table
(pacman_amd/autotune.py) measured fastest (precision "auto"). Beside it, on
the same GPU: the FP64 plain path (one FP64 setup + one-field apply per
field, the reference's calling pattern), the FP64 batched transfer, the
mixed transfer (FP32 factors with FP64 refinement forced on the workspace
patches above 264 points), all five Wendland bases, and the agreement of
the results (rel-L2 vs franke_3d, max differences vs FP64).

--config c2 / c5: the FP64 plain transfer of one field on the jittered-grid
workloads (reference patch heuristic, 150 points per patch).

Prints one JSON line. The CPU baseline is the FP64 oracle (a restatement of
the reference algorithm) timed on this host on a bounded sample of the same
workload: 1 warm-up, median of 3.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector = matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--rbf", type=lambda s: int(s, 0), default=0x11)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-edge", type=float, default=0.5,
                    help="CPU baseline sample: sub-cube [0, edge)^3")
    ap.add_argument("--fields", type=int, default=1,
                    help="fields per transfer (franke_3d plus seeded trig "
                         "fields)")
    ap.add_argument("--mode", default="rbf",
                    choices=["rbf", "apply", "adjoint", "fe-nearest", "mls"],
                    help="rbf: RBF-PUM setup+apply (default); apply: one "
                         "setup, then --steps applies (repeated coupling "
                         "transfer); adjoint: one setup, then forward + "
                         "backward (gradients of 0.5 |u - g|^2 with respect "
                         "to the source values and the target coordinates) "
                         "per step; mls: the moving-least-squares transfer "
                         "(setup + apply) on the same clouds; fe-nearest: the "
                         "nearest-source transfer on the same clouds")
    ap.add_argument("--operator", action="store_true",
                    help="apply mode: store the per-patch operator after the "
                         "setup and apply with it (the build time is "
                         "reported apart)")
    ap.add_argument("--operator-fp32", action="store_true",
                    help="apply mode: as --operator, with the operator stored "
                         "in FP32 (FP64 sums); the max relative difference "
                         "from the FP64 solve is reported")
    ap.add_argument("--graph", action="store_true",
                    help="apply mode: replay the apply as a HIP graph")
    ap.add_argument("--only-headline", action="store_true",
                    help="uneven cloud: time the headline transfer only "
                         "(for profiler runs; the comparisons are null)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB "
                         "(calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported in this build (patch sharding across GPUs is a "
              "design note only).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_probe_mib:
        n = args.copy_probe_mib << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n,
                          "launches": 3}), flush=True)
        return 0

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    log(f"[bench] generating {args.config}")
    if args.config == "c3":
        if args.mode != "rbf":
            print("--config c3 runs the rbf mode only", file=sys.stderr)
            return 2
        return bench_mixture(args)
    S, T, npc, support = W.config(args.config)
    if args.mode == "fe-nearest":
        return bench_fe_nearest(args, S, T)
    if args.mode == "apply":
        return bench_apply(args, S, T, npc, support)
    if args.mode == "adjoint":
        return bench_adjoint(args, S, T, npc, support)
    if args.mode == "mls":
        return bench_mls(args, S, T)
    v = W.franke_3d(*S.T)
    if args.fields > 1:
        v = np.concatenate([v[None], W.trig_fields(S, args.fields - 1,
                                                    seed=11)])
    N, M = len(S), len(T)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dT = torch.from_numpy(T).to(dev)
    dv = torch.from_numpy(v).to(dev)
    torch.cuda.synchronize()

    def step():
        plan = pacman.rbf.Plan(3, args.rbf, dS, dT, npc, 0.15, support)
        out = plan.apply(dv)
        return plan, out

    for i in range(args.warmup):
        plan, out = step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i} done: K={plan.num_patches}")
        del plan, out
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i:
            del plan, out
        plan, out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    log(f"[bench] {args.steps} steps: {ms:.2f} ms/step")

    # Phase times from one more step run in isolation (outside the timed
    # region): in the timed loop a new plan's first reuse of cached device
    # memory waits for the previous step's kernels, which would show up in
    # its first setup phase.
    del plan, out
    torch.cuda.synchronize()
    plan, out = step()
    torch.cuda.synchronize()
    ph = plan.phase_times()
    work = plan.work()
    exact = torch.from_numpy(W.franke_3d(*T.T)).to(dev)
    out0 = out[0] if out.ndim == 2 else out
    rel_l2 = float(torch.linalg.norm(out0 - exact) / torch.linalg.norm(exact))

    # Dominant kernel: the per-patch solve, one launch per size bin (the
    # polynomial fit runs before it and the evaluation after it, each in
    # kernels of their own, timed apart). Algorithmic flops per patch of n
    # sources and m member targets it evaluates itself, counted as for the
    # reference's chain: unpivoted LU (2/3) n^3, forward + back substitution
    # 2 n^2 per field, evaluation 2 n m per field (none: every register-
    # resident patch is evaluated apart); kernel-value evaluations are not
    # counted. The symmetric LDL^T executed here needs n^3 / 3 for the
    # factorisation; that rate is reported as executed_frac.
    Fk = args.fields
    f_alg = (2.0 / 3.0) * work["tile_sum_n3"] + Fk * (
        2 * work["tile_sum_n2"] + 2 * work["tile_sum_nm"])
    f_exec = (1.0 / 3.0) * work["tile_sum_n3"] + Fk * (
        2 * work["tile_sum_n2"] + 2 * work["tile_sum_nm"])
    kern_ms = ph["tiles"]
    achieved = f_alg / (kern_ms / 1e3) / 1e12
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"solve_tile_{args.config}_pmc.json"
                       if args.fields == 1 else
                       f"solve_tile_{args.config}_f{args.fields}_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_apply")

    # Host buffers in and out (the drop-in boundary hands over host arrays):
    # the same transfer with the PCIe copies, outside the timed region.
    # Copies go through the library's pinned staging buffers (allocated by
    # the warm-up call). Median of 3 after 1 warm-up.
    del plan
    torch.cuda.synchronize()

    def host_transfer():
        hp = pacman.rbf.Plan(3, args.rbf, S, T, npc, 0.15, support)
        hp.apply(v)

    host_ms = median_time(host_transfer) * 1e3
    # The drop-in call itself (the reference's fixed patch parameters).
    drop_ms = median_time(lambda: pacman.rbf.interpolate(
        3, 4, args.rbf, S, v if v.ndim == 1 else v[0], T)) * 1e3

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        e = args.cpu_edge
        ms_ = np.all(S < e, axis=1)
        mt_ = np.all(T < e, axis=1)
        Ss, Ts, vs = S[ms_], T[mt_], v[ms_]
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        log(f"[bench] cpu baseline: oracle on {len(Ss)} -> {len(Ts)} "
            f"({cores} threads)")
        t = median_time(lambda: oracle.Setup(3, args.rbf, Ss, Ts, npc, 0.15,
                                             support).apply(vs))
        cpu = {"value": len(Ts) / t / 1e6, "unit": "Mpts/s",
               "cores": cores, "kind": "port",
               "sample": f"{args.config} restricted to [0,{e})^3: "
                         f"{len(Ss)} sources -> {len(Ts)} targets, same "
                         f"parameters, FP64 oracle setup+apply, median of 3 "
                         f"after 1 warm-up: {t:.2f} s"}

    res = {
        "metric": "RBF-PUM setup+apply Mpts/s (FP64 plain path, one field)",
        "value": M / (ms / 1e3) / 1e6,
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: {N} jittered-grid sources -> {M} "
                        f"uniform targets, Wendland 0x{args.rbf:02x}, "
                        f"{npc} nodes per cluster, overlap 0.15, support = "
                        f"patch radius, {args.fields} field(s)",
            "fields": args.fields,
            "mpts_fields_per_s": M * args.fields / (ms / 1e3) / 1e6,
            "patches": int(work["patches"]),
            "max_patch_size": int(work["max_n"]),
            "pairs": int(work["pairs"]),
            "phase_ms": {k: round(x, 3) for k, x in ph.items()},
            "rel_l2_vs_franke_3d": rel_l2,
            "host_buffers_ms": round(host_ms, 3),
            "host_buffers_mpts_s": M / (host_ms / 1e3) / 1e6,
            "drop_in_interpolate_ms": round(drop_ms, 3),
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "solve_tile_kernel (all size-bin launches of one "
                      "apply)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "kernel_ms": kern_ms,
            "flops_per_apply": f_alg,
            "executed_frac": f_exec / (kern_ms / 1e3) / 1e12 /
            PEAK_FP64_TFLOPS,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def median_time(fn, reps=3, warmup=1):
    import time as _t
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = _t.perf_counter()
        fn()
        ts.append(_t.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def bench_mixture(args) -> int:
    """Batched transfer (autotuned precision) vs FP64 plain on the uneven
    (Gaussian-mixture) cloud."""
    import numpy as np
    import torch

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    F = args.fields if args.fields > 1 else 8
    S, T = W.mixture_config()
    V = W.mixture_fields(S, F)
    N, M = len(S), len(T)
    dev = torch.device("cuda:0")
    dS, dT, dV = (torch.from_numpy(a).to(dev) for a in (S, T, V))
    out = torch.empty((F, M), dtype=torch.float64, device=dev)
    leaf, nmin, nmax = 40, 20, 600

    def transfer(rbf, precision, vals, o):
        plan = pacman.rbf.Plan.adaptive(3, rbf, dS, dT, leaf, nmin, nmax,
                                        precision=precision)
        plan.apply(vals, out=o)
        return plan

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, r

    # the headline step: all fields per solve, precision per the size-bin
    # table ("auto")
    ms, plan = timed(lambda: transfer(args.rbf, "auto", dV, out),
                     args.steps, args.warmup)
    log(f"[bench] batched (auto precision): {ms:.2f} ms/step")
    head_out = out.clone()
    work = plan.work()
    large = plan.num_large_patches
    del plan
    # phase times of one more step started on an idle GPU (the host-timed
    # setup phases of back-to-back steps include the previous step's tail)
    torch.cuda.synchronize()
    plan = transfer(args.rbf, "auto", dV, out)
    torch.cuda.synchronize()
    ph = plan.phase_times()
    del plan
    if args.only_headline:
        print(json.dumps({"ms_per_step": ms, "phase_ms": ph, "work": work}),
              flush=True)
        return 0
    # FP64 batched (every class in FP64)
    fp64_out = torch.empty_like(out)
    ms_b, plan = timed(lambda: transfer(args.rbf, "fp64", dV, fp64_out),
                       args.steps, 1)
    del plan
    log(f"[bench] fp64 batched: {ms_b:.2f} ms/step")
    # mixed: FP32 factors + FP64 refinement forced on the workspace path
    mixed_out = torch.empty_like(out)
    ms_m, plan = timed(lambda: transfer(args.rbf, "mixed", dV, mixed_out),
                       args.steps, 1)
    cnt = plan.counters()
    del plan
    log(f"[bench] mixed batched: {ms_m:.2f} ms/step")
    # FP64 plain: one FP64 setup + one-field apply per field
    plain_out = torch.empty_like(out)

    def plain():
        for f in range(F):
            transfer(args.rbf, "fp64", dV[f:f + 1], plain_out[f:f + 1])
        return None

    ms_p, _ = timed(plain, max(1, args.steps // 2), 1)
    log(f"[bench] fp64 plain ({F} transfers): {ms_p:.2f} ms/step")
    exact = torch.from_numpy(W.franke_3d(*T.T)).to(dev)

    def rl2(o):
        return float(torch.linalg.norm(o[0] - exact) / torch.linalg.norm(exact))

    scale = float(fp64_out.abs().max())

    def maxrel(o):
        return float((o - fp64_out).abs().max()) / scale

    def rell2(o, ref):
        return float(torch.linalg.norm(o - ref) / torch.linalg.norm(ref))

    # all five bases: one step each in each precision
    per_basis = {}
    for rbf in (0x10, 0x11, 0x12, 0x13, 0x14):
        o1, o2, o3 = (torch.empty_like(out) for _ in range(3))
        ta, pl = timed(lambda: transfer(rbf, "auto", dV, o1), 1, 1)
        del pl
        tf, pl = timed(lambda: transfer(rbf, "fp64", dV, o2), 1, 1)
        del pl
        tm, pl = timed(lambda: transfer(rbf, "mixed", dV, o3), 1, 1)
        c = pl.counters()
        del pl
        per_basis[f"0x{rbf:02x}"] = {
            "auto_ms": round(ta, 3), "fp64_batched_ms": round(tf, 3),
            "mixed_ms": round(tm, 3),
            "rel_l2_vs_franke_3d": rl2(o1),
            "auto_vs_fp64_rel_l2": rell2(o1, o2),
            "mixed_vs_fp64_rel_l2": rell2(o3, o2),
            "mixed_refinements": c["refinements"],
            "mixed_fallbacks": c["fallbacks"]}
        log(f"[bench] basis 0x{rbf:02x}: auto {ta:.1f} ms, fp64 {tf:.1f} ms, "
            f"mixed {tm:.1f} ms")

    # dominant kernel: the register-resident solve (all size-bin launches of
    # the step; the 8 fields share one launch per bin). Algorithmic flops of
    # the reference's chain per patch: (2/3) n^3 once (one factorisation
    # serves all F fields) + F (2 n^2 + 2 n m), with m the member targets it
    # evaluates itself (the patches with more than 4096 target pairs are
    # evaluated by their own kernel).
    f_tile = (2.0 / 3.0) * work["tile_sum_n3"] + F * (
        2 * work["tile_sum_n2"] + 2 * work["tile_sum_nm"])
    tile_ms = ph["tiles"]
    achieved = f_tile / (tile_ms / 1e3) / 1e12 if tile_ms > 0 else 0.0
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "solve_tile_mixture_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_step")

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        ns = 80000
        Ss, Ts, Vs = S[:ns], T[:ns], V[:, :ns]
        log(f"[bench] cpu baseline: oracle on {ns} -> {ns}, {cores} threads")

        def cpu_run():
            oracle.AdaptiveSetup(3, args.rbf, Ss, Ts, leaf, nmin, nmax).apply(Vs)

        t = median_time(cpu_run)
        cpu = {"value": ns / t / 1e6, "unit": "Mpts/s", "cores": cores,
               "kind": "port",
               "sample": f"the first {ns} of the {N} mixture sources and "
                         f"{M} uniform targets (i.i.d. draws of the same "
                         f"distributions), same builder parameters, {F} "
                         f"fields, FP64 oracle setup+apply (one "
                         f"factorisation per field, as the reference's one "
                         f"field per call); median of 3 after 1 warm-up: "
                         f"{t:.2f} s"}

    res = {
        "metric": "RBF-PUM setup+apply Mpts/s, mixed/batched vs FP64 plain, "
                  "same GPU, equal L2",
        "value": M / (ms / 1e3) / 1e6,
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"uneven cloud: {N} Gaussian-mixture sources (12 "
                        f"components, seed 4) -> {M} uniform targets (seed "
                        f"5), density-aware patches (leaf {leaf}, {nmin} to "
                        f"{nmax} sources, overlap 0.15, support = patch "
                        f"radius), Wendland 0x{args.rbf:02x}, {F} fields, "
                        "all fields per solve, precision per size class "
                        "from the autotuned table (FP64 factors in "
                        "registers up to 264 points; workspace patches in "
                        "the precision the table measured faster)",
            "fields": F,
            "mpts_fields_per_s": M * F / (ms / 1e3) / 1e6,
            "fp64_plain_ms": ms_p,
            "fp64_batched_ms": ms_b,
            "mixed_batched_ms": ms_m,
            "speedup_vs_fp64_plain": ms_p / ms,
            "speedup_vs_fp64_batched": ms_b / ms,
            "speedup_mixed_vs_fp64_plain": ms_p / ms_m,
            "rel_l2_vs_franke_3d": {"batched": rl2(head_out),
                                    "fp64_batched": rl2(fp64_out),
                                    "mixed": rl2(mixed_out),
                                    "fp64_plain": rl2(plain_out)},
            "batched_vs_fp64_max_rel": maxrel(head_out),
            "mixed_vs_fp64_max_rel": maxrel(mixed_out),
            "mixed_vs_fp64_rel_l2": rell2(mixed_out, fp64_out),
            "plain_vs_batched_max_rel": maxrel(plain_out),
            "patches": int(work["patches"]),
            "workspace_patches": large,
            "max_patch_size": int(work["max_n"]),
            "work": work,
            "mixed_refinements": cnt["refinements"],
            "mixed_fallbacks": cnt["fallbacks"],
            "phase_ms": {k: round(x, 3) for k, x in ph.items()},
            "per_basis": per_basis,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "solve_tile_kernel (all size-bin launches of one "
                      "step: FP64 L D L^T on matrix cores, 8 or 16 waves "
                      "per bin from the autotuned table)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "kernel_ms": tile_ms,
            "flops_per_launch": f_tile,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_apply(args, S, T, npc, support) -> int:
    """Repeated apply on one plan (coupling loop): fields change per step."""
    import numpy as np
    import torch

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    F = args.fields
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dT = torch.from_numpy(T).to(dev)
    vals = torch.from_numpy(W.trig_fields(S, F, seed=11)).to(dev)
    out = torch.empty((F, len(T)), dtype=torch.float64, device=dev)
    plan = pacman.rbf.Plan(3, args.rbf, dS, dT, npc, 0.15, support)
    build_ms = fp32_err = None
    use_op = args.operator or args.operator_fp32
    prec = "fp32" if args.operator_fp32 else "fp64"
    if use_op:
        ref = plan.apply(vals) if args.operator_fp32 else None
        torch.cuda.synchronize()
        b0 = time.perf_counter()
        plan.build_operator(precision=prec)
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - b0) * 1e3
        if ref is not None:
            got = plan.apply(vals)
            fp32_err = float((got - ref).abs().max() / ref.abs().max())
            del got, ref
    for _ in range(max(args.warmup, 1)):   # the timed step, untimed
        vals.mul_(-1.0)
        plan.apply(vals, out=out, graph=args.graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        vals.mul_(-1.0)   # new field contents every step, same buffer
        plan.apply(vals, out=out, graph=args.graph)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    M = len(T)
    op_rate = None
    if use_op:
        # whole-step rates of the stored-operator contraction: every entry is
        # read once and used in 2 F flops (FP64 MFMA above 4 fields)
        ob = plan.operator_bytes
        entries = ob / (4 if prec == "fp32" else 8)
        op_rate = {"operator_read_tb_s": ob / (ms / 1e3) / 1e12,
                   "operator_read_frac_of_8tb_s": ob / (ms / 1e3) / 8e12,
                   "contraction_tflop_s": 2 * entries * F / (ms / 1e3) / 1e12,
                   "contraction_frac_of_fp64_peak":
                       2 * entries * F / (ms / 1e3) / 78.6e12}
    res = {
        "metric": "RBF-PUM repeated apply Mpts*fields/s (FP64)",
        "value": M * F / (ms / 1e3) / 1e6, "unit": "Mpts*fields/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: {len(S)} -> {M}, one setup, "
                               f"{F} trig fields per apply, graph="
                               f"{bool(args.graph)}, stored operator="
                               f"{prec if use_op else False}",
                   "fields": F, "graph": bool(args.graph),
                   "operator_build_ms": build_ms,
                   "operator_precision": prec if use_op else None,
                   "operator_fp32_max_rel_vs_fp64": fp32_err,
                   "operator_gb": plan.operator_bytes / 1e9,
                   "operator_rates_per_step": op_rate},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_adjoint(args, S, T, npc, support) -> int:
    """Forward + backward through pacman_amd.autograd on one plan."""
    import numpy as np
    import torch

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dT = torch.from_numpy(T).to(dev)
    f = torch.from_numpy(W.franke_3d(*S.T)).to(dev).requires_grad_()
    g = torch.from_numpy(
        np.random.default_rng(8).normal(size=len(T))).to(dev)
    plan = pacman.rbf.Plan(3, args.rbf, dS, dT, npc, 0.15, support)
    dT.requires_grad_()

    def step():
        f.grad = None
        dT.grad = None
        u = pacman.autograd.transfer(plan, f, dT)
        (0.5 * ((u - g) ** 2).sum()).backward()

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    M = len(T)
    res = {
        "metric": "RBF-PUM forward+backward Mpts/s (FP64)",
        "value": M / (ms / 1e3) / 1e6, "unit": "Mpts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: {len(S)} -> {M}, one setup, "
                               "loss 0.5|u - g|^2, gradients with respect to "
                               "the source values and the target "
                               "coordinates"},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_mls(args, S, T) -> int:
    """Moving-least-squares transfer, inputs in HBM.

    A step = plan setup (grids, k = 10 nearest sources and the QR weights of
    every target) + one apply. CPU baseline: the oracle on 1000 sampled
    targets (its neighbour search is brute force, so it is a port of the
    method, not of a tuned search). Compulsory HBM bytes: sources 24 B read + the
    sorted copy (24 B + 4 B) written, value 8 B read; targets 24 B read, the
    k indices and weights (12 B each) written and read back, 8 B out.
    """
    import torch

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dT = torch.from_numpy(T).to(dev)
    dv = torch.from_numpy(W.franke_3d(*S.T)).to(dev)
    N, M = len(S), len(T)

    def step():
        plan = pacman.mls.Plan(3, dS, dT)
        return plan, plan.apply(dv)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan, out = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # apply alone (the plan kept)
    torch.cuda.synchronize()
    a0 = time.perf_counter()
    for _ in range(10):
        plan.apply(dv)
    torch.cuda.synchronize()
    apply_ms = (time.perf_counter() - a0) * 1e3 / 10
    ref = W.franke_3d(*T.T)
    u = out.cpu().numpy()
    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np
        import oracle
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        sample = T[np.random.default_rng(12).choice(M, 1000, replace=False)]
        fv = W.franke_3d(*S.T)
        c0 = time.perf_counter()
        oracle.mls_interpolate(3, S, fv, sample)
        c1 = time.perf_counter()
        cpu = {"value": len(sample) / (c1 - c0) / 1e6, "unit": "Mpts/s",
               "cores": cores, "kind": "port",
               "sample": f"1000 of the {M} targets against all {N} sources, "
                         f"FP64 oracle (brute-force neighbour search), "
                         f"{c1 - c0:.2f} s"}
    byts = N * (24 + 28 + 8) + M * (24 + 2 * 10 * 12 + 8)
    gbs = byts / (ms / 1e3) / 1e9
    res = {
        "metric": "MLS transfer setup+apply Mpts/s (FP64)",
        "value": M / (ms / 1e3) / 1e6, "unit": "Mpts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config} clouds: {N} jittered-grid "
                               f"sources -> {M} uniform targets, degree-2 "
                               "MLS over the 10 nearest sources",
                   "apply_ms": apply_ms,
                   "rel_l2_vs_franke_3d": float(
                       ((u - ref) ** 2).sum() ** 0.5 / (ref ** 2).sum() ** 0.5)},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "grid build + mls_knn_kernel "
                     "+ mls_weights_kernel + mls_apply_kernel, compulsory "
                     "bytes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_fe_nearest(args, S, T) -> int:
    """Nearest-source transfer (FE NEAREST_NEAREST), inputs in HBM.

    A step = grid build over the sources + one nearest query per target.
    Compulsory HBM bytes: sources 24 B coordinates + 8 B value read, sorted
    copy written (24 B + 4 B index); targets 24 B read, 8 B value + 4 B
    status written.
    """
    import numpy as np
    import torch

    import pacman_amd as pacman
    from pacman_amd import workloads as W

    v = W.franke_3d(*S.T)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dT = torch.from_numpy(T).to(dev)
    dv = torch.from_numpy(v).to(dev)
    N, M = len(S), len(T)
    for _ in range(args.warmup):
        out, st = pacman.fe.nearest_device(3, dS, dv, dT)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, st = pacman.fe.nearest_device(3, dS, dv, dT)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    byts = N * (24 + 8 + 24 + 4) + M * (24 + 8 + 4)
    gbs = byts / (ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        oracle.set_threads(cores)
        c0 = time.perf_counter()
        oracle.fe_nearest(3, S, v, T)
        c1 = time.perf_counter()
        cpu = {"value": M / (c1 - c0) / 1e6, "unit": "Mpts/s",
               "cores": cores, "kind": "port",
               "sample": f"the full workload ({N} -> {M}), FP64 oracle grid "
                         f"search, {c1 - c0:.2f} s"}
    res = {
        "metric": "FE NEAREST_NEAREST transfer Mpts/s",
        "value": M / (ms / 1e3) / 1e6, "unit": "Mpts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config} clouds: {N} jittered-grid "
                               f"sources -> {M} uniform targets, nearest "
                               "source value"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "grid build + "
                     "nearest_value_kernel, compulsory bytes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
