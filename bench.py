"""Benchmark of the FP64 Lax-Friedrichs step on one MI355X.

Workload: the memory-bound scaling case, the converging-diverging channel
r_in = 0.2 + 0.05 sin(pi z), r_out = 0.8 - 0.1 z^2 at 6400 x 2560 (L_max x
M_max; 16,390,400 interior cells), IC rho = 1 (+1e-3 seeded noise), u_z = 0.1,
B_phi = 0.2 / r, p = 0.5, inflow / outflow / slip walls, adaptive CFL 0.4.
A step is one full time step of the loop (stencil, boundary faces, CFL dt
update), with the state resident in HBM.

  value  = GCUPS = interior cells x steps / seconds / 1e9
  roofline.achieved = 128 B (8 FP64 conserved values read + 8 written) x cells
                      / the stencil kernel's average launch time (HIP events)

The CPU baseline times the oracle (the C restatement of the reference's
six-pass step) on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "mpd-plasma-dynamics_amd"), os.path.join(REPO, "tests")]

HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_CELL = 128.0


def workload_config(L, M):
    import cases
    return cases.channel_config(L, M, geometry="condiv", noise=1e-3, max_steps=0)


def cpu_baseline(L, M, steps):
    """The oracle on the host cores: IC setup untimed, `steps` timed steps."""
    import _oracle as O
    import cases
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(cases.oracle_case(workload_config(L, M)))
    o.run(1)  # warm-up step (first touch of the arrays)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    cells = L * (M + 1)
    return {"value": cells * steps / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the full {L}x{M} workload after 1 warm-up step "
                      f"({dt:.1f} s), oracle/lf_oracle.c with OpenMP"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--L", type=int, default=6400)
    ap.add_argument("--M", type=int, default=2560)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU z-slab decomposition is design-only).", file=sys.stderr)
        sys.exit(2)

    import mpd_amd

    L, M = args.L, args.M
    cells = L * (M + 1)
    print(f"bench: {L}x{M} converging-diverging channel, {cells} cells, warmup {args.warmup}, "
          f"steps {args.steps}", file=sys.stderr, flush=True)
    with mpd_amd.Solver(workload_config(L, M)) as sol:
        w0 = time.perf_counter()
        tm = sol.time_steps(args.steps, args.warmup)
        wall = time.perf_counter() - w0
    ms_per_step = tm["total_ms"] / args.steps
    gcups = cells / (ms_per_step * 1e-3) / 1e9
    achieved = ALG_BYTES_PER_CELL * cells / (tm["stencil_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(REPO, "profiles", "stencil_traffic_6400x2560.json")
    if os.path.exists(tf) and (L, M) == (6400, 2560):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": "GCUPS (FP64 LF step, 6400x2560 converging-diverging channel)",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded IC of the workload)",
        "config": {"workload": f"converging-diverging channel {L}x{M}, FP64 LF, adaptive CFL 0.4",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "step_breakdown_ms": {"stencil": round(tm["stencil_ms"], 5), "boundary": round(tm["bc_ms"], 5)},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(L, M, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
