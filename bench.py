"""Benchmark of the FP64 Lax-Friedrichs step on one MI355X.

Default mode (the one JSON line): the memory-bound scaling case, the
converging-diverging channel r_in = 0.2 + 0.05 sin(pi z), r_out = 0.8 - 0.1 z^2
at 6400 x 2560 (L_max x M_max; 16,390,400 interior cells), IC rho = 1 (+1e-3
seeded noise), u_z = 0.1, B_phi = 0.2 / r, p = 0.5, inflow / outflow / slip
walls, adaptive CFL 0.4.  A step is one full time step of the loop (stencil,
boundary faces, CFL dt update), with the state resident in HBM.

  value  = GCUPS = interior cells x steps / seconds / 1e9
  roofline.achieved = 128 B (8 FP64 conserved values read + 8 written)
                      x cells x steps / seconds of the whole timed loop
  kernel_roofline   = the same bytes / the stencil kernel's average launch
                      time (HIP events on the solver's stream)

Also on the line: `batched_sweep` (64 uneven instances at the sweep sizes,
batched against one at a time), `steady_state` (a reduced live run of the
learned warm start at 96x39: generation, training and evaluation) and
`steady_state_400x160` (time to steady state on the warm start's own grid,
400x160, with the committed network trained at that grid; --no-warmstart
skips both, --c3-test 0 the second), the measured copy ceiling of
the HBM, and `cpu_baseline`: the reference's own compiled step
(oracle/_ref/ref_harness, built from the reference sources in the build
container; it travels with the tree) on a bounded sample of the same
workload on this host's physical cores, in a child process pinned as the
reference's job script pins its OpenMP threads (OMP_PROC_BIND=close,
OMP_PLACES=cores), with a one-thread figure beside it, and the oracle (the C
restatement of the reference's six-pass step) timed the same way as `port`.

Other modes (each prints its own report; the profiles' commands use them):
  --mode copy            the copy kernel: counter calibration + copy ceiling
  --mode sweep-probe     a batched sweep workload timed (--probe small|uneven)
  --mode sweep-converge  uneven sweep cases to residual 1e-8, batched vs one at a time
  --mode warmstart       the full learned warm-start experiment
  --mode nn-train        training steps of the warm-start network
  --mode cpu-calibration the oracle against the reference-compiled harness
                         (build container only: needs oracle/_ref)
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "mpd-plasma-dynamics_amd"), os.path.join(REPO, "tests")]

HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_CELL = 128.0
TRAFFIC_PROFILE = os.path.join("profiles", "stencil_traffic_6400x2560.json")


def workload_config(L, M):
    import cases
    return cases.channel_config(L, M, geometry="condiv", noise=1e-3, max_steps=0)


# ------------------------------------------------------------ CPU baseline

def host_topology():
    """Physical cores of the CPUs this process may run on (lscpu's core and
    socket ids), with the machine totals."""
    cpus = sorted(os.sched_getaffinity(0))
    info = {"affinity_cpus": len(cpus), "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu", "-p=cpu,core,socket"], capture_output=True, text=True, check=True).stdout
        rows = [tuple(int(v) for v in ln.split(",")) for ln in out.splitlines() if ln and not ln.startswith("#")]
        allowed = set(cpus)
        cores = {(c, s) for cpu, c, s in rows if cpu in allowed}
        info.update(physical_cores=len(cores), sockets=len({s for _, _, s in rows}),
                    machine_physical_cores=len({(c, s) for _, c, s in rows}))
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in model.splitlines():
            if ln.startswith("Model name:"):
                info["model"] = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError, ValueError):
        info["physical_cores"] = len(cpus)
    return info


def _oracle_child(L, M, steps, threads):
    """Runs the oracle in a child process whose OpenMP runtime starts with
    the given threads, bound close, one place per core (the reference's
    slurm.sh:11-14 settings).  Returns the child's JSON."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--mode", "cpu-oracle-run", "--L", str(L), "--M", str(M),
           "--cpu-steps", str(steps)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def cpu_oracle_run(L, M, steps):
    """(child) the oracle on the workload: IC setup and one warm-up step
    untimed, `steps` timed steps."""
    import _oracle as O
    import cases
    o = O.Oracle(cases.oracle_case(workload_config(L, M)))
    o.run(1)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    print(json.dumps({"seconds": dt, "steps": steps, "gcups": L * (M + 1) * steps / dt / 1e9,
                      "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}), flush=True)


REF_HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")


def _ref_child(L, M, steps, threads):
    """The reference's own compiled step (oracle/_ref/ref_harness, timing
    mode) with OpenMP pinned as slurm.sh:11-14 pins it."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = subprocess.run([REF_HARNESS, "time", str(L), str(M), str(steps)], env=env, capture_output=True,
                         text=True, check=True).stdout
    kv = dict(ln.split(None, 1) for ln in out.splitlines() if ln.strip())
    secs = float(kv["loop_seconds"])
    return {"seconds": secs, "steps": steps, "gcups": L * (M + 1) * steps / secs / 1e9}


def cpu_baseline(L, M, steps, steps_1t, ref_steps, ref_steps_1t):
    """The reference's own step (when oracle/_ref/ref_harness was built with
    the tree) and the oracle port, each with all physical cores of this
    process (at most 16) and with one thread."""
    topo = host_topology()
    threads = max(1, min(topo.get("physical_cores", 1), 16))
    pin = "OMP_PROC_BIND=close OMP_PLACES=cores"
    port = {}
    many = _oracle_child(L, M, steps, threads)
    port = {"value": many["gcups"], "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the full {L}x{M} workload after 1 untimed step ({many['seconds']:.1f} s), "
                      f"oracle/lf_oracle.c (six passes per step, as src/solver.cpp:22-37) with OpenMP, "
                      f"{threads} threads, {pin}"}
    if steps_1t > 0:
        one = _oracle_child(L, M, steps_1t, 1)
        port["one_thread"] = {"value": one["gcups"], "unit": "GCUPS",
                              "sample": f"{steps_1t} steps after 1 untimed step ({one['seconds']:.1f} s)"}
    if not os.path.exists(REF_HARNESS):
        port["host"] = topo
        return port
    ref = _ref_child(L, M, ref_steps, threads)
    out = {"value": ref["gcups"], "unit": "GCUPS", "cores": threads, "kind": "reference",
           "sample": f"{ref_steps} steps of the {L}x{M} workload after 1 untimed step ({ref['seconds']:.1f} s): "
                     f"the reference's Solver::Advance + Diagnostics::ComputeDt compiled from its own sources "
                     f"(oracle/_ref/ref_harness, timing mode; no boundary segments, whose O(L+M) work per step "
                     f"is left out), {threads} threads, {pin}",
           "host": topo, "port": port}
    if ref_steps_1t > 0:
        one = _ref_child(L, M, ref_steps_1t, 1)
        out["one_thread"] = {"value": one["gcups"], "unit": "GCUPS",
                             "sample": f"{ref_steps_1t} steps after 1 untimed step ({one['seconds']:.1f} s)"}
    return out


def cpu_calibration():
    """The oracle against the reference's own compiled solver (oracle/_ref/
    ref_harness, no boundary segments: the state the harness can run) on the
    same state and threads: loop wall time of each, ratio oracle / reference."""
    import numpy as np
    import _oracle as O
    if not O.ref_harness_available():
        sys.exit("oracle/_ref/ref_harness is missing (build it with make -C oracle ref where the reference is)")
    topo = host_topology()
    threads = int(os.environ.get("OMP_NUM_THREADS", topo.get("physical_cores", 1)))
    ic = {"rho": "1.0", "v_z": "0.1", "v_r": "0.0", "v_phi": "0.0", "H_z": "0.0", "H_r": "0.0",
          "H_phi": "0.2 / r", "e": "0.5 / ((gamma - 1) * rho)"}
    rows = []
    for L, M, steps, reps in ((200, 80, 10, 5), (1600, 640, 10, 3)):
        gp = [0.2, 0.05, 0.8, 0.1]
        case = dict(L=L, M=M, gamma=1.6666666666666667, geom=(O.GEOM_CONDIV, gp), ic=ic, cfl=0.4, growth=1.1,
                    dt_min=1e-12, dt_max=1e-2, dt=1e-4, adaptive=True)
        o0 = O.Oracle(case)
        phys = [o0.phys(k).copy() for k in range(10)]
        u0 = [o0.u0(k).copy() for k in range(8)]
        ref_t, orc_t = [], []
        for _ in range(reps):
            out = O.run_ref_harness(L, M, O.GEOM_CONDIV, gp, 1.6666666666666667, steps, 1e-4, 1, 0.4, 1.1, 1e-12,
                                    1e-2, phys, u0)
            ref_t.append(out["loop_seconds"])
            o = O.Oracle(case)
            for k in range(10):
                o.phys(k)[:] = phys[k]
            for k in range(8):
                o.u0(k)[:] = u0[k]
            dt = 1e-4
            t0 = time.perf_counter()
            for _s in range(steps):
                o.advance(dt)
                dt = o.compute_dt(dt)
            orc_t.append(time.perf_counter() - t0)
            same = all(np.array_equal(np.nan_to_num(o.u0(k)), np.nan_to_num(out["u0"][k])) for k in range(8))
        r, c = min(ref_t), min(orc_t)
        rows.append({"grid": f"{L}x{M}", "steps": steps, "reps": reps, "reference_s": r, "oracle_s": c,
                     "oracle_over_reference": c / r, "same_state": bool(same)})
    print(json.dumps({"threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                      "omp_places": os.environ.get("OMP_PLACES"), "host": topo, "runs": rows}, indent=1))


# ------------------------------------------------------------- GPU extras

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
    steady states, then time plain cold solves (8 timed, scaled) and cold
    batched solves of `test` held-out cases to residual 1e-8, and warm
    batched solves to 1e-9 (the warm start's stop rule); errors of both
    against the cases' cold solves continued to residual 1e-12."""
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
                          budget=WS.step_budget(steps[conv]), warm_levels=[1e-9], ref_threshold=1e-12)
    keep = ["plain_cold_total_s", "cold_batch_s", "warm_total_s", "speedup_vs_plain_cold", "speedup_vs_cold_batch",
            "steps_ratio_cold_over_warm", "fallbacks", "all_converged_warm", "warm_finish_p50_s",
            "warm_finish_p99_s", "cold_finish_p50_s", "cold_finish_p99_s", "warm_threshold",
            "max_err_cold_vs_reference", "max_err_warm_vs_reference", "max_field_err_warm_vs_reference"]
    out = {k: r[k] for k in keep if k in r}
    out["reference_states_reached"] = r.get("reference", {}).get("reached")
    out.update({"workload": f"{test} held-out sweep cases at {L}x{M}: cold to residual 1e-8, warm to 1e-9, "
                            f"references to 1e-12; network trained on {int(conv.sum())} steady states, "
                            f"{epochs} epochs",
                "generate_s": round(gen_s, 2), "train_s": round(train_s, 2), "final_loss": losses[-1]})
    return out


WARMSTART_NET = os.path.join("mpd-plasma-dynamics_amd", "data", "warmstart_400x160.npz")


def warmstart_c3_measure(test, plain, ref_steps, L=400, M=160):
    """Time to steady state on the warm start's own grid (400x160) with the
    committed network (trained by the project's kernels on 2000 solver steady
    states at this grid; profiles/warmstart_400x160_training.json): `test`
    held-out cases (the first of the 64 of the recorded evaluation), plain
    cold solves one at a time (`plain` timed, scaled), cold batched to 1e-8,
    warm batched to 1e-9, and both against the cases' cold solves continued
    to 1e-12 (at most `ref_steps` more steps; the cases that reach it)."""
    import warmstart as WS
    import sweep
    path = os.path.join(REPO, WARMSTART_NET)
    net, norm, train_steps = WS.load_network(path, L, M)
    r, _, _ = WS.evaluate(net, norm, sweep.sample_params(test, 2025), L, M, plain_sample=plain,
                          budget=WS.step_budget(train_steps), warm_levels=[1e-9], ref_threshold=1e-12,
                          ref_max_steps=ref_steps, max_steps=600000)
    keep = ["plain_cold_total_s", "cold_batch_s", "warm_total_s", "speedup_vs_plain_cold", "speedup_vs_cold_batch",
            "steps_ratio_cold_over_warm", "fallbacks", "all_converged_warm", "warm_finish_p50_s",
            "warm_finish_p99_s", "cold_finish_p50_s", "cold_finish_p99_s", "warm_threshold",
            "max_err_cold_vs_reference", "max_err_warm_vs_reference", "max_field_err_warm_vs_reference"]
    out = {k: r[k] for k in keep if k in r}
    out["reference_states_reached"] = r.get("reference", {}).get("reached")
    out["workload"] = (f"{test} held-out sweep cases at {L}x{M} (config 3 grid): plain cold solves one at a time "
                       f"({plain} timed, scaled), cold batched to residual 1e-8, warm batched to 1e-9, references "
                       f"to 1e-12 (at most {ref_steps} more steps); every case capped at 600000 steps; "
                       f"network {WARMSTART_NET}")
    return out


def traffic_record():
    """HBM bytes per stencil launch from the committed counter profile (two
    rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, scaled by the copy
    kernel's calibration), with its source named."""
    path = os.path.join(REPO, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("bytes_per_launch"), {"file": TRAFFIC_PROFILE, "measured": rec.get("measured", ""),
                                         "kernel": rec.get("kernel", "")}


# ------------------------------------------------------------------ modes

def mode_step(args):
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
    alg = ALG_BYTES_PER_CELL * cells
    achieved = alg / (ms_per_step * 1e-3) / 1e9
    k_achieved = alg / (tm["stencil_ms"] * 1e-3) / 1e9
    traffic, tsrc = traffic_record() if (L, M) == (6400, 2560) else (None, None)
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "definition": "128 B x cells x steps / wall time of the timed steps (whole step)",
                     "traffic_source": tsrc},
        "kernel_roofline": {"kernel": "lf_stencil_tiled", "achieved": round(k_achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(k_achieved / HBM_PEAK_GBS, 4),
                            "launch_ms": round(tm["stencil_ms"], 5)},
        "step_breakdown_ms": {"stencil": round(tm["stencil_ms"], 5), "boundary": round(tm["bc_ms"], 5)},
        "wall_s": round(wall, 3),
    }
    for k in ("tile_order", "tile_rows"):
        if getattr(args, k) is not None:
            line["config"][k] = getattr(args, k)
    if not args.no_copy:
        c = mpd_amd.bench_copy(1024, 0)
        line["hbm_copy_ceiling"] = {"gbs": round(c["ceiling_gbs"], 1), "frac_of_peak":
                                    round(c["ceiling_gbs"] / HBM_PEAK_GBS, 4),
                                    "shape": f"{c['best_blocks']} workgroups, {c['best_stores']} stores, 1 GiB buffers"}
    if not args.no_sweep:
        print("bench: batched sweep", file=sys.stderr, flush=True)
        line["batched_sweep"] = sweep_measure()
    if not args.no_warmstart:
        print("bench: learned warm start (reduced live run)", file=sys.stderr, flush=True)
        line["steady_state"] = warmstart_measure(args.warmstart_train, args.warmstart_test, args.warmstart_epochs)
        if args.c3_test > 0:
            print("bench: learned warm start at 400x160 (committed network)", file=sys.stderr, flush=True)
            try:
                line["steady_state_400x160"] = warmstart_c3_measure(args.c3_test, args.c3_plain, args.c3_ref_steps)
            except Exception as e:  # the headline line is printed whatever happens here
                line["steady_state_400x160"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(L, M, args.cpu_steps, args.cpu_steps_1t, args.ref_steps,
                                            args.ref_steps_1t)
    print(json.dumps(line), flush=True)


def mode_copy(args):
    import mpd_amd
    c = mpd_amd.bench_copy(args.copy_mib, 3)
    print(json.dumps(c), flush=True)


def mode_sweep_probe(args):
    import mpd_amd
    import sweep
    with mpd_amd.Batch(sweep.probe_configs(args.probe)) as b:
        b.run(args.warmup)
        print(b.autotune(5), flush=True)
        if args.persistent:
            b.set_option("persistent", 1)
        t0 = time.perf_counter()
        b.run(args.steps)
        el = time.perf_counter() - t0
    mode = "persistent launches" if args.persistent else "graphs of per-phase launches"
    print(f"{args.probe} ({mode}): {args.steps} steps {el:.3f} s = {el / args.steps * 1e6:.1f} us/step", flush=True)


def mode_sweep_converge(args):
    import mpd_amd
    import sweep
    r = sweep.converge_uneven(mpd_amd, [int(v) for v in args.sizes.split(",")])
    print(json.dumps(r, indent=1), flush=True)


def mode_warmstart(args):
    import warmstart
    levels = [float(v) for v in args.warm_levels.split(",")] if args.warm_levels else None
    warmstart.experiment(args.L, args.M, args.train, args.test, args.epochs, args.batch, args.lr, args.clip,
                         args.plain_sample, args.seed, args.out, warm_levels=levels,
                         ref_threshold=args.ref_threshold, load=args.load, gen_batch=args.gen_batch,
                         net_grid=tuple(int(v) for v in args.net_grid.split("x")) if args.net_grid else None,
                         ref_max_steps=args.ref_max_steps, case_max_steps=args.case_max_steps,
                         staged=args.warm_staged)


def mode_nn_train(args):
    import surrogate
    ms = surrogate.time_training(args.H, args.W, args.batch, args.steps)
    print(f"training step, batch {args.batch} at {args.H}x{args.W}: {ms:.2f} ms", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="step",
                    choices=["step", "copy", "sweep-probe", "sweep-converge", "warmstart", "nn-train",
                             "cpu-calibration", "cpu-oracle-run"])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--L", type=int, default=None)
    ap.add_argument("--M", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-steps-1t", type=int, default=4)
    ap.add_argument("--ref-steps", type=int, default=40)
    ap.add_argument("--ref-steps-1t", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--tile-order", type=int, default=None,
                    help="stencil tile order (library default 0); for profiling the alternatives")
    ap.add_argument("--tile-rows", type=int, default=None, help="stencil tile height (library default 8)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--warmstart-train", type=int, default=500)
    ap.add_argument("--warmstart-test", type=int, default=64)
    ap.add_argument("--warmstart-epochs", type=int, default=60)
    ap.add_argument("--no-warmstart", action="store_true")
    ap.add_argument("--c3-test", type=int, default=32,
                    help="held-out cases of the 400x160 warm-start measurement (0: skip it)")
    ap.add_argument("--c3-plain", type=int, default=4, help="plain cold solves timed there")
    ap.add_argument("--c3-ref-steps", type=int, default=200000, help="reference continuation cap there")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--probe", choices=["small", "uneven"], default="small")
    ap.add_argument("--persistent", action="store_true",
                    help="sweep-probe: run the batch as persistent launches (one workgroup per instance)")
    ap.add_argument("--sizes", default="48,96,144,192")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--clip", type=float, default=None)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default="")
    ap.add_argument("--warm-levels", default="", help="tighter warm-start stop levels, e.g. 1e-9,1e-10")
    ap.add_argument("--warm-staged", action="store_true",
                    help="run the whole warm batch to each level in turn and compare the states of every level "
                         "(default: each warm case runs straight to the last level)")
    ap.add_argument("--ref-threshold", type=float, default=None,
                    help="also build reference steady states converged to this level and report errors against them")
    ap.add_argument("--load", default="", help="a saved warmstart_net.npz (skips generation and training)")
    ap.add_argument("--gen-batch", type=int, default=512, help="cases per batch of the steady-state generation")
    ap.add_argument("--ref-max-steps", type=int, default=2000000,
                    help="most steps a reference state runs beyond the sweep threshold")
    ap.add_argument("--case-max-steps", type=int, default=2000000, help="time.max_steps of every evaluated case")
    ap.add_argument("--net-grid", default="", help="grid LxM of a loaded network when it differs from --L/--M "
                    "(its prediction is interpolated onto the run's grid)")
    ap.add_argument("--H", type=int, default=96)
    ap.add_argument("--W", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU z-slab decomposition is design-only).", file=sys.stderr)
        sys.exit(2)
    m = args.mode
    if m in ("step", "cpu-oracle-run"):
        args.L = args.L or 6400
        args.M = args.M or 2560
    if m == "step":
        args.steps = args.steps or 10000
        mode_step(args)
    elif m == "cpu-oracle-run":
        cpu_oracle_run(args.L, args.M, args.cpu_steps)
    elif m == "cpu-calibration":
        cpu_calibration()
    elif m == "copy":
        mode_copy(args)
    elif m == "sweep-probe":
        args.steps = args.steps or 2000
        mode_sweep_probe(args)
    elif m == "sweep-converge":
        mode_sweep_converge(args)
    elif m == "warmstart":
        args.L = args.L or 96
        args.M = args.M or 39
        mode_warmstart(args)
    elif m == "nn-train":
        args.steps = args.steps or 40
        mode_nn_train(args)


if __name__ == "__main__":
    main()
