"""Benchmark of the FP64 Lax-Friedrichs step on one MI355X.

Also on the line: `batched_sweep` (64 uneven instances at the sweep sizes, batched
against one at a time) and `steady_state` (a reduced live run of the learned warm
start: wall time to residual 1e-8 of surrogate-warm-started batched solves against
cold-started plain solves; --no-warmstart skips it).

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


def sweep_measure(steps=1000, warm=200, n=64, seed=12345):
    """Batched sweep at the uneven sizes nz in {200, 400, 800, 1600}, nr = 0.4 nz
    (n seeded inflow cases): per-step time of the batch, against the sum of
    the per-step times of single solves of each size (one case at a time)."""
    import numpy as np
    import mpd_amd
    import sweep
    rng = np.random.default_rng(seed)
    sizes, cfgs = [], []
    for _ in range(n):
        nz = int(rng.choice([200, 400, 800, 1600]))
        q = [rng.uniform(0.5, 2.0), rng.uniform(0.05, 0.3), rng.uniform(0.05, 0.5), 0.05, 0.1]
        sizes.append(nz)
        cfgs.append(sweep.case_config(q, nz, int(nz * 0.4), threshold=0.0, max_steps=0))
    cells = sum(nz * (int(nz * 0.4) + 1) for nz in sizes)
    with mpd_amd.Batch(cfgs) as b:
        b.run(warm)
        b.autotune(5)
        t0 = time.perf_counter()
        b.run(steps)
        batch_s = (time.perf_counter() - t0) / steps
    single = {}
    for nz in sorted(set(sizes)):
        with mpd_amd.Solver(cfgs[sizes.index(nz)]) as sol:
            sol.run(warm)
            t0 = time.perf_counter()
            sol.run(steps)
            single[nz] = (time.perf_counter() - t0) / steps
    seq_s = sum(single[nz] for nz in sizes)
    return {"workload": f"{n} instances, nz in {{200,400,800,1600}} (counts "
                        f"{[sizes.count(z) for z in (200, 400, 800, 1600)]}), nr = 0.4 nz, {cells} cells",
            "batched_ms_per_step": round(batch_s * 1e3, 4), "batched_gcups": round(cells / batch_s / 1e9, 3),
            "one_at_a_time_ms_per_step": round(seq_s * 1e3, 4), "one_at_a_time_gcups": round(cells / seq_s / 1e9, 3),
            "speedup": round(seq_s / batch_s, 2), "steps_timed": steps}


def warmstart_measure(train, test, epochs, L=96, M=39):
    """Reduced live run of the learned warm start: train on `train` solver
    steady states, then time plain cold solves (8 timed, scaled), cold and
    warm batched solves of `test` held-out cases to residual 1e-8."""
    import warmstart as WS
    import surrogate as SG
    import sweep
    import numpy as np
    t0 = time.perf_counter()
    params, fields, steps, conv = WS.generate(train, L, M, seed=2024, max_steps=300000, batch=512)
    gen_s = time.perf_counter() - t0
    params, fields = params[conv], fields[conv]
    radii = WS.grid_radii(params, L, M)
    base = SG.baseline_fields(params, radii, L, M, WS.GAMMA)
    norm = WS.Normalizer(fields - base)
    net = SG.UNet(L, M + 1, seed=2024)
    t1 = time.perf_counter()
    losses = WS.train(net, SG.features(params, radii, L, M), norm.encode(fields - base).astype(np.float32), epochs,
                      32, 2e-3, log_every=max(1, epochs // 4))
    train_s = time.perf_counter() - t1
    r, _, _ = WS.evaluate(net, norm, sweep.sample_params(test, 2025), L, M, plain_sample=8,
                          budget=WS.step_budget(steps[conv]))
    keep = ["plain_cold_total_s", "cold_batch_s", "warm_total_s", "speedup_vs_plain_cold", "speedup_vs_cold_batch",
            "steps_ratio_cold_over_warm", "fallbacks", "all_converged_warm", "warm_finish_p50_s",
            "warm_finish_p99_s", "max_field_err_warm_vs_cold"]
    out = {k: r[k] for k in keep}
    out.update({"workload": f"{test} held-out sweep cases at {L}x{M}, residual 1e-8; network trained on "
                            f"{int(conv.sum())} steady states, {epochs} epochs",
                "generate_s": round(gen_s, 2), "train_s": round(train_s, 2), "final_loss": losses[-1]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--L", type=int, default=6400)
    ap.add_argument("--M", type=int, default=2560)
    ap.add_argument("--cpu-steps", type=int, default=80)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-order", type=int, default=None,
                    help="stencil tile order (library default 0); for profiling the alternatives")
    ap.add_argument("--tile-rows", type=int, default=None, help="stencil tile height (library default 8)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--warmstart-train", type=int, default=500)
    ap.add_argument("--warmstart-test", type=int, default=64)
    ap.add_argument("--warmstart-epochs", type=int, default=60)
    ap.add_argument("--no-warmstart", action="store_true")
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
        if args.tile_order is not None:
            sol.set_option("tile_order", args.tile_order)
        if args.tile_rows is not None:
            sol.set_option("tile_rows", args.tile_rows)
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
    for k in ("tile_order", "tile_rows"):
        if getattr(args, k) is not None:
            line["config"][k] = getattr(args, k)
    if not args.no_sweep:
        print("bench: batched sweep", file=sys.stderr, flush=True)
        line["batched_sweep"] = sweep_measure()
    if not args.no_warmstart:
        print("bench: learned warm start (reduced live run)", file=sys.stderr, flush=True)
        line["steady_state"] = warmstart_measure(args.warmstart_train, args.warmstart_test, args.warmstart_epochs)
    if not args.no_cpu:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(L, M, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
