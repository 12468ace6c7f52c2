"""Parameter sweeps of the LF solver: config texts of seeded sweep cases and
the batched runs that take them to a steady state.

A sweep case is the converging-diverging coaxial channel
r_in(z) = 0.2 + a_in sin(pi z), r_out(z) = 0.8 - a_out z^2 with a fixed
inflow at z = 0 (rho_in, u_z = u_in, B_phi = amp / r, p = 0.5), zero-gradient
outflow at z = 1 and slip walls, started from the uniform inflow state (the
cold start), gamma = 5/3, adaptive CFL 0.4.  Its 5 parameters are
(rho_in, u_in, amp, a_in, a_out).  The text is config.yaml in the
reference's schema (include/config.hpp:119-177) with the schema extensions
documented for this library (time.max_steps, geometry converging_diverging).
"""
import time

import numpy as np

PARAM_NAMES = ["rho_in", "u_in", "amp", "a_in", "a_out"]
PARAM_LO = np.array([0.5, 0.05, 0.05, 0.0, 0.0])
PARAM_HI = np.array([2.0, 0.3, 0.5, 0.1, 0.1])


def sample_params(n, seed, fixed_walls=None):
    """n seeded uniform draws over the parameter box (SplitMix-free: numpy's
    PCG64 with the given seed).  fixed_walls=(a_in, a_out) pins the wall
    amplitudes."""
    rng = np.random.default_rng(seed)
    p = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random((n, 5))
    if fixed_walls is not None:
        p[:, 3], p[:, 4] = fixed_walls
    return p


def case_config(params, L_max, M_max, threshold=1e-8, check_frequency=100, max_steps=0, dt=1.0e-4):
    rho, uz, amp, a_in, a_out = (repr(float(v)) for v in params)
    return f"""physics:
  gamma: 1.6666666666666667
  beta: 0.05
  H_z0: 0.25
time:
  T: 1.0e9
  dt: {dt}
  max_steps: {int(max_steps)}
adaptive_dt:
  enabled: true
  cfl_number: 0.4
  growth_factor: 1.1
  dt_min: 1.0e-12
  dt_max: 1.0e-2
convergence:
  threshold: {threshold}
  check_frequency: {int(check_frequency)}
grid:
  L_max: {int(L_max)}
  M_max: {int(M_max)}
geometry:
  type: converging_diverging
  params: {{ r_inner: 0.2, a_inner: {a_in}, r_outer: 0.8, a_outer: {a_out} }}
initial_conditions:
  rho: {rho}
  v_z: {uz}
  v_r: 0.0
  v_phi: 0.0
  H_z: 0.0
  H_r: 0.0
  H_phi: "{amp} / r"
  e: "0.5 / ((gamma - 1) * rho)"
boundary_conditions:
  l_lo:
    - rho: {{ dirichlet: {rho} }}
      v_z: {{ dirichlet: {uz} }}
      v_r: {{ dirichlet: 0.0 }}
      v_phi: {{ dirichlet: 0.0 }}
      e: "0.5 / ((gamma - 1) * rho)"
      H_z: {{ dirichlet: 0.0 }}
      H_r: {{ dirichlet: 0.0 }}
      H_phi: "{amp} / r"
  l_hi:
    - rho: neumann
      v_z: neumann
      v_r: neumann
      v_phi: neumann
      e: neumann
      H_z: neumann
      H_r: neumann
      H_phi: neumann
  m_lo:
    - rho: neumann
      v_z: neumann
      v_r: "v_z_nb * r_z_nb"
      v_phi: neumann
      e: neumann
      H_z: neumann
      H_r: "H_z_nb * r_z_nb"
      H_phi: neumann
  m_hi:
    - rho: neumann
      v_z: neumann
      v_r: "v_z * r_z"
      v_phi: neumann
      e: neumann
      H_z: neumann
      H_r: "H_z * r_z"
      H_phi: neumann
"""


def run_batch(mpd, configs, warm_u0=None, max_steps=0, autotune=True):
    """Run the cases of `configs` together to their own ends (convergence,
    step cap).  warm_u0: optional list (per case, None for a cold start) of 8
    conserved arrays to start from.  Returns (results, wall seconds, batch):
    the batch stays open so the caller can read the states."""
    b = mpd.Batch(configs)
    if warm_u0 is not None:
        for i, u in enumerate(warm_u0):
            if u is not None:
                b.warm_start(i, u)
    if autotune:
        b.autotune(5)
    t0 = time.perf_counter()
    res = b.run(max_steps)
    return res, time.perf_counter() - t0, b


def probe_configs(kind, seed=7):
    """The profiling workloads: 'small' = 256 sweep cases at 96x39, 'uneven' =
    64 cases with nz in {200, 400, 800, 1600}, nr = 0.4 nz (cold starts,
    convergence checks every 100 steps at threshold 1e-8)."""
    if kind == "small":
        return [case_config(q, 96, 39) for q in sample_params(256, seed)]
    rng = np.random.default_rng(12345)
    cfgs = []
    for _ in range(64):
        nz = int(rng.choice([200, 400, 800, 1600]))
        q = [rng.uniform(0.5, 2.0), rng.uniform(0.05, 0.3), rng.uniform(0.05, 0.5), 0.05, 0.1]
        cfgs.append(case_config(q, nz, int(nz * 0.4)))
    return cfgs


def converge_uneven(mpd_amd, sizes, n=64, seed=2, threshold=1e-8, plain_sample=16, chunk=2000):
    """n seeded cases with nz drawn uniformly from `sizes` (nr = 0.4 nz, the
    parameter box of sample_params), every case run to SolutionChange <
    threshold or to 3 M steps: once batched, once one case at a time
    (plain_sample cases timed, scaled to n by the mean of each size).
    Returns a dict (with how many cases converged each way)."""
    import json  # noqa: F401  (the caller prints)
    rng = np.random.default_rng(seed)
    nz = rng.choice(np.asarray(sizes), size=n)
    params = sample_params(n, seed)
    cfgs = [case_config(params[i], int(nz[i]), max(2, int(round(0.4 * nz[i]))), threshold=threshold,
                        max_steps=3000000) for i in range(n)]
    cells = sum(int(z) * (max(2, int(round(0.4 * z))) + 1) for z in nz)
    finish = np.full(n, np.nan)
    steps = np.zeros(n, int)
    sizes_of = np.array([int(z) * (max(2, int(round(0.4 * z))) + 1) for z in nz])
    with mpd_amd.Batch(cfgs) as b:
        b.autotune(5)
        t0 = time.perf_counter()
        last = 0.0
        tuned_cells = cells
        retunes = 0
        while True:
            res = b.run(chunk)
            el = time.perf_counter() - t0
            steps += np.array([r["steps"] for r in res])
            for i, r in enumerate(res):
                if r["done"] and np.isnan(finish[i]):
                    finish[i] = el
            nd = sum(r["done"] for r in res)
            # When the running cases hold under a quarter of the cells the
            # tile heights were tuned for, tune them again (inside the timed
            # run): the tail of a few large cases wants shorter tiles.
            live_cells = int(sum(sizes_of[i] for i, r in enumerate(res) if not r["done"]))
            if 0 < live_cells < tuned_cells // 4:
                b.autotune(3)
                tuned_cells = live_cells
                retunes += 1
            if el - last >= 10.0 or nd == n:
                print(f"[sweep] batched: {el:.1f} s, {nd}/{n} done, max steps {steps.max()}", flush=True)
                last = el
            if nd == n:
                break
        batch_s = time.perf_counter() - t0
        conv = all(r["converged"] for r in res)
    # One case at a time: the same configs, sampled evenly over the sizes.
    per_size = {}
    plain_conv = 0
    order = np.argsort(nz, kind="stable")
    pick = order[np.linspace(0, n - 1, min(n, plain_sample)).astype(int)]
    for i in pick:
        with mpd_amd.Solver(cfgs[i]) as sol:
            t1 = time.perf_counter()
            r = sol.run()
            per_size.setdefault(int(nz[i]), []).append(time.perf_counter() - t1)
            plain_conv += int(r["converged"])
        print(f"[sweep] one case at a time: nz {nz[i]}, {r['steps']} steps, {per_size[int(nz[i])][-1]:.2f} s, "
              f"converged {r['converged']}", flush=True)
    mean = {z: float(np.mean(v)) for z, v in per_size.items()}
    plain_s = float(sum(mean[int(z)] for z in nz))
    return {"cases": n, "sizes": [int(z) for z in sizes], "counts": [int(np.sum(nz == z)) for z in sizes],
            "cells": int(cells), "threshold": threshold, "all_converged": bool(conv),
            "converged_batched": int(sum(r["converged"] for r in res)), "step_cap": 3000000,
            "batched_s": batch_s, "one_at_a_time_s": plain_s, "one_at_a_time_timed": int(len(pick)),
            "one_at_a_time_converged": plain_conv,
            "one_at_a_time_mean_s_by_size": mean, "speedup": plain_s / batch_s,
            "steps_p50": float(np.median(steps)), "steps_max": int(steps.max()),
            "finish_p50_s": float(np.nanmedian(finish)), "finish_p99_s": float(np.nanpercentile(finish, 99)),
            "retunes": retunes}
