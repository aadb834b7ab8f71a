"""Size-bin autotuner of the per-patch solve (run on the GPU).

Builds one density-aware plan on the uneven-cloud workload (4,096,000
Gaussian-mixture sources -> 4,096,000 uniform targets, the bench's default
transfer), then, for every field count F = 1 .. 8 solved per launch, times
each variant of each class of the per-patch solve on the plan's own patches
(median of `--reps` launches, HIP events, after the value gather and the
polynomial fit):

* register-resident bins R = ceil((n + F) / 16): 8 waves per workgroup (two
  workgroups per CU, bins 1 .. 13) and 16 waves (one per CU, bins
  11 .. 17); where both exist the faster one is kept;
* the workspace path (patches above 264 points): FP64 factors, or FP32
  factors with FP64 refinement (plus its FP64 fallback pass); precision
  "auto" uses the faster one.

Writes the table as a C++ header (pacman_amd/csrc/rbf_tune_table.hpp by
default; rebuild the library afterwards) and the measured times as JSON.

    python -m pacman_amd.autotune --out <header> --report <json>
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_R = 17
FMAX = 8


def variants(R):
    """Wave counts with an instantiated kernel for bin R."""
    v = []
    if R <= 13:
        v.append(8)
    if R >= 11:
        v.append(16)
    return v


def tune(reps=5, n=4_096_000, rbf=0x11, log=print):
    import torch

    from pacman_amd import rbf as rbf_mod
    from pacman_amd import workloads as W

    dev = torch.device("cuda:0")
    S, T = W.mixture_config(n)
    V = W.mixture_fields(S, FMAX)
    dS, dT, dV = (torch.from_numpy(a).to(dev) for a in (S, T, V))
    plan = rbf_mod.Plan.adaptive(3, rbf, dS, dT, 40, 20, 600)
    log(f"[autotune] plan: K={plan.num_patches}, "
        f"{plan.num_large_patches} workspace patches, "
        f"largest {plan.max_patch_size}")
    times = {}
    waves = [[8 if R <= 10 else 16 for R in range(MAX_R + 1)]
             for _ in range(FMAX)]
    ws_fp32 = [False] * FMAX
    for F in range(1, FMAX + 1):
        vals = dV[:F].contiguous()
        tf = {"bins": {}, "workspace": {}}
        for R in range(1, MAX_R + 1):
            row = {}
            for w in variants(R):
                ms = plan.time_solve(vals, R, w, reps=reps)
                if ms >= 0:
                    row[w] = ms
            if row:
                tf["bins"][R] = row
                waves[F - 1][R] = min(row, key=row.get)
        for prec in ("fp64", "mixed"):
            ms = plan.time_solve(vals, 0, ws_precision=prec, reps=reps)
            if ms >= 0:
                tf["workspace"][prec] = ms
        wsd = tf["workspace"]
        ws_fp32[F - 1] = ("mixed" in wsd and "fp64" in wsd
                          and wsd["mixed"] < wsd["fp64"])
        times[F] = tf
        log(f"[autotune] F={F}: " + ", ".join(
            f"R{R}:" + "/".join(f"{w}w {ms:.2f}" for w, ms in row.items())
            for R, row in tf["bins"].items()) +
            f"; workspace {wsd}")
    info = {
        "device": f"{torch.cuda.get_device_name(0)} "
                  f"({torch.cuda.get_device_properties(0).gcnArchName})",
        "workload": f"{n} Gaussian-mixture sources (seed 4) -> {n} uniform "
                    f"targets (seed 5), density-aware patches (leaf 40, 20 "
                    f"to 600 sources), Wendland 0x{rbf:02x}",
        "reps": reps,
        "date": time.strftime("%Y-%m-%d"),
    }
    return waves, ws_fp32, times, info


def header(waves, ws_fp32, times, info) -> str:
    lines = [
        "// Size-bin table of the per-patch solve, generated by "
        "pacman_amd/autotune.py",
        f"// on {info['device']} ({info['date']}), workload: "
        f"{info['workload']};",
        f"// median of {info['reps']} launches per variant. Times in ms "
        "(F fields per launch):",
    ]
    for F in range(1, FMAX + 1):
        tf = times.get(F, {"bins": {}, "workspace": {}})
        ch = []
        for R, row in tf["bins"].items():
            if len(row) > 1:
                ch.append(f"R{R} " + "/".join(
                    f"{w}w {ms:.2f}" for w, ms in row.items()))
        ws = tf["workspace"]
        ch.append("workspace " + "/".join(
            f"{k} {v:.2f}" for k, v in ws.items()))
        lines.append(f"//   F={F}: " + "; ".join(ch))
    lines += [
        "#pragma once",
        "",
        "namespace pacman_amd {",
        "",
        "// waves[F - 1][R]: workgroup waves of the register-resident solve of "
        "bin R",
        "// with F fields per launch; wsFp32[F - 1]: precision \"auto\" "
        "factors the",
        "// workspace-path patches in FP32 with FP64 refinement.",
        "struct TuneTable {",
        f"  int waves[{FMAX}][{MAX_R + 1}];",
        f"  bool wsFp32[{FMAX}];",
        "};",
        "",
        "constexpr TuneTable kTune = {",
        "    {",
    ]
    for F in range(FMAX):
        lines.append("        {" + ", ".join(str(w) for w in waves[F]) + "},")
    lines += [
        "    },",
        "    {" + ", ".join("true" if b else "false" for b in ws_fp32) + "},",
        "};",
        "",
        "} // namespace pacman_amd",
        "",
    ]
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc",
                                                  "rbf_tune_table.hpp"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--points", type=int, default=4_096_000)
    args = ap.parse_args(argv)
    waves, ws_fp32, times, info = tune(
        args.reps, args.points,
        log=lambda *a: print(*a, file=sys.stderr, flush=True))
    with open(args.out, "w") as fh:
        fh.write(header(waves, ws_fp32, times, info))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump({"info": info, "waves": waves, "ws_fp32": ws_fp32,
                       "times_ms": times}, fh, indent=1)
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
