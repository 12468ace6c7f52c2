"""The learned warm-start pipeline: steady states from the batched solver,
training of the encoder-decoder on the GPU, prediction of initial states, and
the warm-versus-cold comparison in which the plain LF loop is both the arbiter
(each warm start is relaxed to the same convergence threshold as a cold
start) and the fallback (a case whose residual stalls starts again cold).

The full experiment is `experiment` (bench.py --mode warmstart).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import mpd_amd  # noqa: E402
import surrogate as SG  # noqa: E402
import sweep  # noqa: E402

GAMMA = 1.6666666666666667


def progress(msg):
    print(f"[warmstart {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def grid_radii(params, L, M):
    """(N, L+2, M+3) cell radii of each case's grid: Grid::Build's
    r = (1 - frac) r_in(z) + frac r_out(z), z = (l - 1) / L, frac = (m - 1) / M
    (src/grid.cpp:24-37), ghost ring included."""
    params = np.asarray(params, np.float64)
    z = (np.arange(L + 2) - 1.0) * (1.0 / L)
    frac = (np.arange(M + 3) - 1.0) * (1.0 / M)
    r_in = 0.2 + params[:, 3, None] * np.sin(np.pi * z)[None, :]
    r_out = 0.8 - params[:, 4, None] * (z * z)[None, :]
    return (1.0 - frac)[None, None, :] * r_in[:, :, None] + frac[None, None, :] * r_out[:, :, None]


def solve_batch(configs, warm_u0=None, chunk=0, label="batch"):
    """Runs a batch to the end of every case's loop; with chunk > 0 in slices
    of chunk steps, recording when (wall seconds from the start) each case
    finished.  Returns (results, wall seconds, finish times, batch)."""
    b = mpd_amd.Batch(configs)
    if warm_u0 is not None:
        for i, u in enumerate(warm_u0):
            if u is not None:
                b.warm_start(i, u)
    b.autotune(5)
    n = len(configs)
    cells = [(lm[0]) * (lm[1] + 1) for lm in (mpd_amd.config_check(c) for c in configs)]
    tuned_cells = sum(cells)
    steps = np.zeros(n, int)
    finish = np.full(n, np.nan)
    t0 = time.perf_counter()
    last = -1e9
    while True:
        res = b.run(chunk)
        el = time.perf_counter() - t0
        steps += np.array([r["steps"] for r in res])
        for i, r in enumerate(res):
            if r["done"] and np.isnan(finish[i]):
                finish[i] = el
        ndone = sum(r["done"] for r in res)
        # The running cases' tile heights tuned again (inside the timed run)
        # when they hold under a quarter of the cells of the last tuning.
        live = sum(c for c, r in zip(cells, res) if not r["done"])
        if chunk and 0 < live < tuned_cells // 4:
            b.autotune(3)
            tuned_cells = live
        if chunk and (el - last >= 10.0 or ndone == n):
            progress(f"{label}: {el:.1f}s, {ndone}/{n} done, max steps {steps.max()}")
            last = el
        if ndone == n or not chunk:
            break
    for i, r in enumerate(res):
        r["steps"] = int(steps[i])
    return res, time.perf_counter() - t0, finish, b


def steady_fields(b, n, L, M):
    out = np.empty((n, L, M + 1, 8))
    for i in range(n):
        _, ph = b.instance(i).download(u0=False)
        out[i] = SG.fields_from_phys(ph, L, M)
    return out


def generate(n, L, M, seed, threshold=1e-8, max_steps=400000, batch=512):
    """n seeded sweep cases solved cold to the threshold by the batched
    solver, up to `batch` cases per batch (each case's result is the same
    whatever it is batched with; one large batch pays the latency-bound tail
    of its slowest cases once).  Returns (params, fields (n, L, Mn, 8), steps,
    converged)."""
    params = sweep.sample_params(n, seed)
    fields = np.empty((n, L, M + 1, 8))
    steps = np.zeros(n, int)
    conv = np.zeros(n, bool)
    for s0 in range(0, n, batch):
        idx = list(range(s0, min(n, s0 + batch)))
        cfgs = [sweep.case_config(params[i], L, M, threshold=threshold, max_steps=max_steps) for i in idx]
        res, wall, _, b = solve_batch(cfgs, chunk=20000, label=f"generate {s0}-{idx[-1]}")
        fields[idx] = steady_fields(b, len(idx), L, M)
        steps[idx] = [r["steps"] for r in res]
        conv[idx] = [bool(r["converged"]) for r in res]
        b.close()
    return params, fields, steps, conv


class Normalizer:
    """Targets: (steady fields - uniform inflow state - mean) / std per
    channel, statistics over the training set.  A channel without variance
    in the training set (a field the sweep never excites, e.g. v_phi, H_z,
    H_r here) is predicted as its mean exactly."""

    def __init__(self, resid, mean=None, std=None):
        self.mean = resid.reshape(-1, 8).mean(axis=0) if mean is None else np.asarray(mean, np.float64)
        self.std = resid.reshape(-1, 8).std(axis=0) if std is None else np.asarray(std, np.float64)
        self.live = self.std > 1e-12 * np.maximum(1.0, np.abs(self.mean))

    def encode(self, resid):
        return np.where(self.live, (resid - self.mean) / np.where(self.live, self.std, 1.0), 0.0)

    def decode(self, out):
        return np.where(self.live, out * self.std, 0.0) + self.mean

    def to_json(self):
        return {"mean": self.mean.tolist(), "std": self.std.tolist()}


def train(net, X, T, epochs, batch, lr, seed=0, log_every=20, clip=None, guard=3.0, stats=None):
    """Adam on the GPU kernels.  X: (N, H, W, 8) float32 features, T: (N, H,
    W, 8) float32 normalized targets (host arrays, uploaded once).  Cosine
    learning-rate decay.  Divergence guard: the weights of the best epoch so
    far are kept; an epoch whose mean loss exceeds guard x the best restores
    them, restarts Adam and halves the learning rate (a late gradient spike
    against Adam's small second moments can otherwise leave the network
    stuck at a constant output).  clip: optional global gradient-norm clip.
    Returns the per-epoch mean losses; `stats` (a dict) receives the number
    of clipped steps, the largest gradient norm and the rollbacks."""
    X = SG.pad_grid(X, net.H, net.W)
    T = SG.pad_grid(T, net.H, net.W)
    N, H, W, C = X.shape
    Xd = SG.DeviceArray.from_numpy(X.reshape(N, -1))
    Td = SG.DeviceArray.from_numpy(T.reshape(N, -1))
    xb = SG.DeviceArray((batch, H, W, C))
    xb16 = SG.DeviceArray((batch, H, W, C), "bf16")
    tb = SG.DeviceArray((batch, H, W, T.shape[-1]))
    dy = SG.DeviceArray((batch, H, W, T.shape[-1]))
    idx = SG.DeviceArray(batch, "i32")
    L = SG.nn_lib()
    rng = np.random.default_rng(seed)
    nb = N // batch
    total = epochs * nb
    losses = []
    it = 0
    clipped, gmax = 0, 0.0
    best, best_state, lr_scale, rollbacks = np.inf, None, 1.0, 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(N)
        acc = 0.0
        for j in range(nb):
            idx.upload(perm[j * batch:(j + 1) * batch].astype(np.int32))
            SG._ck(L.mpd_nn_gather_rows(SG._p(Xd), SG._p(idx), SG._p(xb), batch, H * W * C, None))
            SG._ck(L.mpd_nn_to_bf16(SG._p(xb), SG._p(xb16), xb.size, None))
            SG._ck(L.mpd_nn_gather_rows(SG._p(Td), SG._p(idx), SG._p(tb), batch, tb.size // batch, None))
            y = net.forward(xb16, batch)
            acc += net.mse(y, tb, dy)
            net.backward(dy, batch)
            lr_t = 0.5 * lr * lr_scale * (1 + np.cos(np.pi * it / total))
            g = net.adam_step(float(lr_t), clip=clip)
            if g is not None:
                gmax = max(gmax, g)
                clipped += int(g > clip)
            it += 1
        losses.append(acc / nb)
        if guard and best_state is not None and losses[-1] > guard * best:
            net.restore(best_state)
            lr_scale *= 0.5
            rollbacks += 1
            progress(f"epoch {ep}: loss {losses[-1]:.4g} > {guard} x best {best:.4g}: back to the best weights, "
                     f"learning rate x {lr_scale}")
        elif losses[-1] < best:
            best = losses[-1]
            if guard:
                best_state = net.snapshot()
        if ep % log_every == 0 or ep == epochs - 1:
            extra = f", {clipped} clipped steps, largest gradient norm {gmax:.3g}" if clip is not None else ""
            progress(f"epoch {ep}: loss {losses[-1]:.4g} ({time.perf_counter() - t0:.1f}s{extra})")
    if guard and best_state is not None and losses[-1] > guard * best:
        net.restore(best_state)
    if stats is not None:
        stats.update(clipped_steps=clipped, max_grad_norm=gmax, steps=it, rollbacks=rollbacks)
    return losses


def predict(net, norm, params, radii, L, M, batch=32):
    """Predicted steady physical fields (n, L, Mn, 8) of the cases."""
    n = len(params)
    X = SG.features(params, radii, L, M)
    base = SG.baseline_fields(params, radii, L, M, GAMMA)
    X = SG.pad_grid(X, net.H, net.W)
    out = np.empty((n, L, M + 1, 8))
    xb16 = SG.DeviceArray((batch, net.H, net.W, SG.N_FEATURES), "bf16")
    for s0 in range(0, n, batch):
        k = min(batch, n - s0)
        xx = np.zeros((batch,) + X.shape[1:], np.float32)
        xx[:k] = X[s0:s0 + k]
        xb16.upload(xx)
        y = net.forward(xb16, batch).download()[:k, :L, :M + 1]
        out[s0:s0 + k] = base[s0:s0 + k] + norm.decode(y.astype(np.float64))
    return out


def resample_fields(f, L0, M0, L1, M1):
    """Fields (n, L0, M0 + 1, 8) on one grid of the unit (z, frac) square
    (cell l at z = (l - 1) / L0, node m at frac = (m - 1) / M0) bilinearly
    interpolated onto another, constant beyond the source's last row."""
    f = np.asarray(f, np.float64)
    z0 = np.arange(L0) / L0
    z1 = np.arange(L1) / L1
    y0 = np.arange(M0 + 1) / M0
    y1 = np.arange(M1 + 1) / M1
    iz = np.clip(np.searchsorted(z0, z1, side="right") - 1, 0, L0 - 2)
    wz = np.clip((z1 - z0[iz]) / (z0[iz + 1] - z0[iz]), 0.0, 1.0)
    iy = np.clip(np.searchsorted(y0, y1, side="right") - 1, 0, M0 - 1)
    wy = np.clip((y1 - y0[iy]) / (y0[iy + 1] - y0[iy]), 0.0, 1.0)
    rows = f[:, iz] * (1.0 - wz)[None, :, None, None] + f[:, iz + 1] * wz[None, :, None, None]
    return rows[:, :, iy] * (1.0 - wy)[None, None, :, None] + rows[:, :, iy + 1] * wy[None, None, :, None]


def predict_on_grid(net, norm, params, L_net, M_net, L, M, batch=32):
    """The network's prediction on its own grid (L_net x M_net), interpolated
    onto an L x M grid of the same cases."""
    radii = grid_radii(params, L_net, M_net)
    f = predict(net, norm, params, radii, L_net, M_net, batch=batch)
    return f if (L_net, M_net) == (L, M) else resample_fields(f, L_net, M_net, L, M)


def warm_u0(fields, radii, L, M):
    rows, cols = L + 2, M + 3
    return [SG.u0_from_fields(fields[i], radii[i], rows, cols) for i in range(len(fields))]


def field_errors(a, b):
    """Per-field relative L-inf error max|a - b| / max|b| (OUT_FIELDS order;
    fields that are identically zero in b report the absolute error)."""
    errs = []
    for k in range(8):
        scale = np.max(np.abs(b[..., k]))
        d = np.max(np.abs(a[..., k] - b[..., k]))
        errs.append(float(d / scale) if scale > 0 else float(d))
    return errs


def admissible(fields):
    """A predicted state the solver may start from: finite everywhere, with
    positive density and specific internal energy (so positive pressure)."""
    f = np.asarray(fields)
    return bool(np.all(np.isfinite(f)) and np.all(f[..., 0] > 0) and np.all(f[..., 4] > 0))


def stalled(done, converged, change, change_mid, stall_change):
    """Fallback rule after the stall check's steps: a case that ended without
    converging (a non-finite state ends the loop once t becomes NaN), or that
    still runs with a non-finite SolutionChange, or whose SolutionChange is
    above stall_change and did not fall over the second half of the check's
    steps."""
    if done:
        return not converged
    if not np.isfinite(change):
        return True
    return bool(change > stall_change and not (change < change_mid))


def fallback_rule(change, least, steps, grow, budget):
    """The warm run's fallback tests at a chunk end for one running case:
    returns (the case's new smallest SolutionChange, rising: the residual is
    not finite or above grow x that smallest value, over: past the step
    budget).  grow / budget None: that test is off."""
    if np.isfinite(change) and change > 0.0:
        least = min(least, change)
    rising = grow is not None and (not np.isfinite(change) or change > grow * least)
    over = budget is not None and steps > budget
    return least, bool(rising), bool(over)


def step_budget(train_steps, factor=1.5):
    """Most steps a warm-started case may take before it is restarted cold:
    factor times the slowest cold solve among the training states (a warm
    start that needs more than any cold start did has gone wrong)."""
    train_steps = np.asarray(train_steps)
    return int(factor * train_steps.max()) if train_steps.size else None


def set_threshold(b, n, level, skip=()):
    """A new convergence stop level for every instance of batch b but those
    in `skip`."""
    for i in range(n):
        if i not in skip:
            b.instance(i).set_option("convergence_threshold", level)


def run_to_end(b, n, chunk, steps, finish=None, t_start=None, on_chunk=None, label="batch"):
    """Runs batch b until every instance's loop has ended, in slices of
    `chunk` steps; adds the steps taken to `steps` and records in `finish`
    (seconds since t_start) when each instance ended.  on_chunk(res) may
    restart instances (it returns the set it restarted).  Returns the last
    results and the seconds spent in the runs."""
    spent = 0.0
    last = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        res = b.run(chunk)
        spent += time.perf_counter() - t0
        steps += np.array([r["steps"] for r in res])
        if time.perf_counter() - last >= 10.0:
            last = time.perf_counter()
            progress(f"{label}: {spent:.1f}s, {sum(r['done'] for r in res)}/{n} done, max steps {steps.max()}")
        if finish is not None:
            el = time.perf_counter() - t_start
            for i, r in enumerate(res):
                if r["done"] and np.isnan(finish[i]):
                    finish[i] = el
        if on_chunk is not None:
            for i in on_chunk(res):
                if finish is not None:
                    finish[i] = np.nan
        if all(r["done"] for r in res):
            return res, spent


def reference_states(cfgs, L, M, ref_threshold, max_steps, chunk=20000):
    """Tightly converged FP64 steady states: each case's cold solve (bit-exact
    with the oracle's) continued from its 1e-8 state to SolutionChange <
    ref_threshold, or to max_steps more steps (stagnation at the rounding
    floor).  Returns (fields, steps, reached)."""
    n = len(cfgs)
    b = mpd_amd.Batch(cfgs)
    b.autotune(5)
    steps = np.zeros(n, int)
    run_to_end(b, n, chunk, steps, label="reference (cold to the sweep threshold)")
    set_threshold(b, n, ref_threshold)
    cap = np.zeros(n, int)
    t0 = last = time.perf_counter()
    while True:
        res = b.run(chunk)
        for i, r in enumerate(res):
            steps[i] += r["steps"]
            cap[i] += r["steps"]
        if time.perf_counter() - last >= 10.0:
            last = time.perf_counter()
            progress(f"reference to {ref_threshold:g}: {last - t0:.1f}s, {sum(r['done'] for r in res)}/{n} done, "
                     f"max steps {steps.max()}")
        if all(r["done"] or cap[i] >= max_steps for i, r in enumerate(res)):
            break
    reached = np.array([bool(r["converged"]) for r in res])
    fields = steady_fields(b, n, L, M)
    b.close()
    return fields, steps, reached


def evaluate(net, norm, params, L, M, threshold=1e-8, stall_steps=2000, stall_change=1e-6, plain_sample=None,
             chunk=1000, fallback=True, pred=None, budget=None, warm_levels=None, ref_threshold=None,
             ref_max_steps=2000000, max_steps=2000000, staged=False, grow=10.0):
    """Cold-started plain solves (each case alone, mpd_run) against the
    surrogate-warm-started batched sweep, both to the same threshold.  A
    warm case whose SolutionChange is still above stall_change after
    stall_steps steps is restarted cold (fallback), and so is one still
    running after `budget` steps (step_budget; None: no budget), and one
    whose SolutionChange rises above `grow` times the smallest value it has
    reached (a warm state that excites a growing mode: its residual falls,
    then climbs for good; None: no such rule).
    plain_sample: time only that many plain solves and scale to all cases
    (stated in the result).
    warm_levels: tighter convergence levels for the warm-started cases (the
    warm start's stop rule is the last one; the wall time and steps reported
    are to it).  staged=False: every warm case runs straight to the last
    level (the timing protocol: each case stops when it reaches it).
    staged=True: the whole batch runs to `threshold`, then to each level in
    turn, and the states at every level are compared (the accuracy study; a
    case's finish time then waits for the slowest case of the level before).
    Cases that start or restart cold always stop at `threshold`.
    ref_threshold: also build tightly converged FP64 steady states
    (reference_states) and report every state's per-field error against
    them."""
    n = len(params)
    cfgs = [sweep.case_config(p, L, M, threshold=threshold, max_steps=max_steps) for p in params]
    radii = grid_radii(params, L, M)
    levels = [threshold] + [v for v in (warm_levels or []) if v < threshold]
    if not staged:
        levels = levels[-1:]

    # Plain cold solves, one case at a time (the reference's way: one run per case).
    m = n if plain_sample is None else min(n, plain_sample)
    plain_steps, plain_t = [], []
    for i in range(m):
        with mpd_amd.Solver(cfgs[i]) as s:
            t0 = time.perf_counter()
            r = s.run()
            plain_t.append(time.perf_counter() - t0)
            plain_steps.append(r["steps"])
        progress(f"plain cold solve {i + 1}/{m}: {r['steps']} steps, {plain_t[-1]:.2f}s")
    plain_total = float(np.sum(plain_t)) * (n / m)
    progress(f"plain cold solves: {m} timed, {np.mean(plain_t):.3f}s mean, {plain_total:.2f}s for {n}")

    # Cold batched (the batching gain alone) and its steady states.
    cold_res, cold_wall, cold_finish, bc = solve_batch(cfgs, chunk=chunk, label="cold batch")
    cold_fields = steady_fields(bc, n, L, M)
    bc.close()
    progress(f"cold batch: {cold_wall:.2f}s")

    # Warm batched: inference, warm start, stall check, fallback.
    t0 = time.perf_counter()
    if pred is None:
        pred = predict(net, norm, params, radii, L, M)
    elif callable(pred):
        pred = pred()  # e.g. a prediction on a coarser grid, interpolated (inside the timing)
    infer = time.perf_counter() - t0
    b = mpd_amd.Batch(cfgs)
    rejected = [i for i in range(n) if not admissible(pred[i])]
    for i, u in enumerate(warm_u0(pred, radii, L, M)):
        if i not in rejected:
            b.warm_start(i, u)
    if levels[0] != threshold:
        set_threshold(b, n, levels[0], skip=set(rejected))
    b.autotune(5)
    t1 = time.perf_counter()
    # Stall check: after stall_steps, a case still running with its
    # SolutionChange above stall_change that did not decrease over the second
    # half of those steps (or is not finite) starts again cold.
    res = b.run(stall_steps // 2)
    steps = np.array([r["steps"] for r in res])
    mid = np.array([r["change"] for r in res])
    res = b.run(stall_steps - stall_steps // 2)
    steps += np.array([r["steps"] for r in res])
    relax = time.perf_counter() - t1
    finish = np.full(n, np.nan)
    for i, r in enumerate(res):
        if r["done"]:
            finish[i] = relax
    stall = [i for i, r in enumerate(res) if fallback and i not in rejected and
             stalled(r["done"], r["converged"], r["change"], mid[i], stall_change)]
    for i in stall:
        b.instance(i).cold_restart(cfgs[i])
        b.instance(i).set_option("convergence_threshold", threshold)
        finish[i] = np.nan
    restarted = set(stall) | set(rejected)
    late = []

    least = np.full(n, np.inf)  # smallest SolutionChange each warm case has reached
    rising = []

    def over_budget(res):
        # Step budget: a warm case still running past it starts again cold;
        # so does one whose residual has risen grow-fold above its minimum.
        out = []
        if not fallback:
            return out
        for i, r in enumerate(res):
            if r["done"] or i in restarted:
                continue
            least[i], up, over = fallback_rule(r["change"], least[i], steps[i], grow, budget)
            if up or over:
                b.instance(i).cold_restart(cfgs[i])
                b.instance(i).set_option("convergence_threshold", threshold)
                restarted.add(i)
                (rising if up else late).append(i)
                out.append(i)
        return out

    level_fields, level_steps, level_wall = [], [], []
    res, spent = run_to_end(b, n, chunk, steps, finish, t1, over_budget, label=f"warm batch to {levels[0]:g}")
    relax += spent
    for k, level in enumerate(levels):
        if k > 0:
            # Per-case finish times are those of the last level.  A case
            # restarted cold keeps the sweep threshold (a cold solve's own
            # stop rule); the step budget still applies to the warm ones.
            set_threshold(b, n, level, skip=restarted)
            finish[:] = np.nan
            res, spent = run_to_end(b, n, chunk, steps, finish, t1, over_budget, label=f"warm batch to {level:g}")
            relax += spent
        level_steps.append(steps.copy())
        level_wall.append(relax)
        level_fields.append(steady_fields(b, n, L, M))
    warm_wall = relax
    warm_fields = level_fields[-1]
    b.close()
    progress(f"warm batch: inference {infer:.3f}s, relax {warm_wall:.2f}s, {len(rejected)} rejected, "
             f"{len(stall)} stalled, {len(late)} over the step budget, {len(rising)} with a rising residual")

    errs = np.array([field_errors(warm_fields[i], cold_fields[i]) for i in range(n)])
    pred_errs = np.array([field_errors(pred[i], cold_fields[i]) for i in range(n)])
    cold_steps = np.array([r["steps"] for r in cold_res])
    out = {
        "cases": n, "grid": [L, M], "threshold": threshold, "warm_threshold": levels[-1],
        "plain_cold_total_s": plain_total, "plain_timed_cases": m,
        "plain_cold_steps_mean": float(np.mean(plain_steps)),
        "cold_batch_s": cold_wall, "cold_steps_total": int(cold_steps.sum()),
        "cold_steps_p50": float(np.median(cold_steps)),
        "cold_steps_max": int(cold_steps.max()),
        "warm_inference_s": infer, "warm_batch_s": warm_wall, "warm_total_s": infer + warm_wall,
        "warm_steps_p50": float(np.median(steps)), "warm_steps_max": int(steps.max()),
        "steps_saved_mean": float(np.mean(cold_steps - steps)),
        "steps_ratio_cold_over_warm": float(np.sum(cold_steps) / np.sum(steps)),
        "fallbacks": len(rejected) + len(stall) + len(late) + len(rising), "rejected_predictions": len(rejected),
        "stalled": len(stall), "over_budget": len(late), "rising": len(rising), "step_budget": budget,
        "all_converged_cold": bool(all(r["converged"] for r in cold_res)),
        "all_converged_warm": bool(all(r["converged"] for r in res)),
        "speedup_vs_plain_cold": plain_total / (infer + warm_wall),
        "speedup_vs_cold_batch": cold_wall / (infer + warm_wall),
        "warm_finish_p50_s": float(np.nanmedian(finish)), "warm_finish_p99_s": float(np.nanpercentile(finish, 99)),
        "cold_finish_p50_s": float(np.nanmedian(cold_finish)),
        "cold_finish_p99_s": float(np.nanpercentile(cold_finish, 99)),
        "max_field_err_warm_vs_cold": dict(zip(SG.OUT_FIELDS, errs.max(axis=0).tolist())),
        "max_field_err_prediction_vs_cold": dict(zip(SG.OUT_FIELDS, pred_errs.max(axis=0).tolist())),
        "median_field_err_prediction_vs_cold": dict(zip(SG.OUT_FIELDS, np.median(pred_errs, axis=0).tolist())),
        "warm_levels": [{"threshold": lv, "steps_total": int(level_steps[k].sum()),
                         "steps_p50": float(np.median(level_steps[k])), "relax_s": level_wall[k]}
                        for k, lv in enumerate(levels)],
    }
    if ref_threshold is not None:
        t2 = time.perf_counter()
        ref_fields, ref_steps, reached = reference_states(cfgs, L, M, ref_threshold, ref_max_steps)
        progress(f"reference states: {reached.sum()}/{n} reached {ref_threshold:g} "
                 f"({time.perf_counter() - t2:.1f}s, steps p50 {np.median(ref_steps):.0f})")

        def worst(states):
            # Over the cases whose reference reached the level (a case that
            # stagnates short of it has no converged reference to measure
            # against); per-case worst field for the distribution.
            e = np.array([field_errors(states[i], ref_fields[i]) for i in range(n)])
            per_case = e.max(axis=1)
            ok = e[reached] if reached.any() else e
            return (dict(zip(SG.OUT_FIELDS, ok.max(axis=0).tolist())), float(ok.max()),
                    {"p50": float(np.median(per_case[reached])) if reached.any() else None,
                     "p90": float(np.percentile(per_case[reached], 90)) if reached.any() else None,
                     "cases_le_1e-6": int(np.sum(per_case[reached] <= 1e-6)),
                     "per_case_all": per_case.tolist()})

        cold_ref, cold_max, cold_dist = worst(cold_fields)
        out.update({"reference": {"threshold": ref_threshold, "reached": int(reached.sum()),
                                  "reached_per_case": reached.astype(int).tolist(),
                                  "steps_p50": float(np.median(ref_steps)), "steps_max": int(ref_steps.max()),
                                  "max_steps_beyond_threshold": ref_max_steps},
                    "max_field_err_cold_vs_reference": cold_ref, "max_err_cold_vs_reference": cold_max,
                    "err_cold_vs_reference_dist": cold_dist})
        for k, lv in enumerate(levels):
            e, mx, dist = worst(level_fields[k])
            out["warm_levels"][k].update({"max_field_err_vs_reference": e, "max_err_vs_reference": mx,
                                          "err_vs_reference_dist": dist})
        out["max_field_err_warm_vs_reference"] = out["warm_levels"][-1]["max_field_err_vs_reference"]
        out["max_err_warm_vs_reference"] = out["warm_levels"][-1]["max_err_vs_reference"]
    out["per_case"] = {"cold_steps": [int(r["steps"]) for r in cold_res], "warm_steps": steps.astype(int).tolist(),
                       "warm_converged": [int(r["converged"]) for r in res],
                       "restarted_cold": sorted(int(i) for i in restarted),
                       "cold_finish_s": np.nan_to_num(cold_finish, nan=-1.0).tolist(),
                       "warm_finish_s": np.nan_to_num(finish, nan=-1.0).tolist()}
    return out, warm_fields, cold_fields


def load_network(path, L, M, seed=2024):
    """A network and normalizer saved by experiment(out=...) (its
    warmstart_net.npz, FP32 weights w<k>) or by save_network_bf16 (bf16 weight
    bit patterns wb<k>), with the training cases' cold step counts."""
    z = np.load(path)
    net = SG.UNet(L, M + 1, seed=seed)
    nl = len(net.layers)
    if "wb0" in z:
        ws = [(z[f"wb{k}"].astype(np.uint32) << 16).view(np.float32) for k in range(nl)]
    else:
        ws = [z[f"w{k}"] for k in range(nl)]
    net.load_params(ws, [z[f"b{k}"] for k in range(nl)])
    norm = Normalizer(None, mean=z["mean"], std=z["std"])
    return net, norm, z["train_steps"] if "train_steps" in z else np.array([])


def save_network_bf16(src, dst):
    """Rewrites a warmstart_net.npz with its weights as the bf16 bit patterns
    the kernels compute with (round to nearest even of the FP32 masters, the
    rounding of the weights' bf16 copies), biases, normalizer and training
    step counts as they are: half the size, the same predictions."""
    z = np.load(src)
    nl = sum(1 for k in z.files if k.startswith("w") and k[1:].isdigit())
    out = {f"wb{k}": SG.bf16_bits(z[f"w{k}"]) for k in range(nl)}
    out.update({f"b{k}": z[f"b{k}"].astype(np.float32) for k in range(nl)})
    out.update(mean=z["mean"], std=z["std"], train_steps=z["train_steps"])
    np.savez_compressed(dst, **out)


def experiment(L, M, train_n=2000, test_n=64, epochs=150, batch=32, lr=2e-3, clip=None, plain_sample=16,
               seed=2024, out="", warm_levels=None, ref_threshold=None, load="", gen_batch=512, net_grid=None,
               ref_max_steps=2000000, case_max_steps=2000000, staged=False):
    """The whole experiment: train_n steady states generated, the network
    trained on them (or a saved one loaded), test_n held-out cases evaluated
    (evaluate); prints the result as JSON and, with `out`, writes it and the
    trained weights there."""
    Ln, Mn_ = net_grid if net_grid else (L, M)
    if load:
        net, norm, train_steps = load_network(load, Ln, Mn_, seed)
        result_extra = {"loaded_network": os.path.basename(load), "network_grid": [Ln, Mn_]}
        progress(f"network loaded from {load} (grid {Ln}x{Mn_})")
    else:
        t0 = time.perf_counter()
        params, fields, steps, conv = generate(train_n, L, M, seed, batch=gen_batch)
        gen_s = time.perf_counter() - t0
        keep = conv
        progress(f"generated {train_n} steady states in {gen_s:.1f}s ({keep.sum()} converged, steps p50 "
                 f"{np.median(steps):.0f} max {steps.max()})")
        params, fields = params[keep], fields[keep]
        train_steps = steps[keep]
        radii = grid_radii(params, L, M)
        base = SG.baseline_fields(params, radii, L, M, GAMMA)
        norm = Normalizer(fields - base)
        X = SG.features(params, radii, L, M)
        T = norm.encode(fields - base).astype(np.float32)
        net = SG.UNet(L, M + 1, seed=seed)
        progress(f"network: {net.n_params()} parameters, {len(net.layers)} convs")
        t1 = time.perf_counter()
        tstats = {}
        losses = train(net, X, T, epochs, batch, lr, seed=seed, clip=clip, stats=tstats)
        train_s = time.perf_counter() - t1
        result_extra = {"train_cases": int(keep.sum()), "generate_s": gen_s, "train_s": train_s, "epochs": epochs,
                        "clip": clip, "rollbacks": tstats["rollbacks"], "final_loss": losses[-1],
                        "first_loss": losses[0], "train_steps_p50": float(np.median(train_steps)),
                        "train_steps_max": int(train_steps.max())}
        if out:
            os.makedirs(out, exist_ok=True)
            sd = net.state_dict()
            np.savez(os.path.join(out, "warmstart_net.npz"), **{f"w{k}": w for k, w in enumerate(sd["w"])},
                     **{f"b{k}": b for k, b in enumerate(sd["b"])}, mean=norm.mean, std=norm.std,
                     train_steps=train_steps)
    if test_n <= 0:
        # Training only (the network is in `out` for a later evaluation).
        result_extra["n_params"] = net.n_params()
        print(json.dumps(result_extra, indent=1), flush=True)
        if out:
            with open(os.path.join(out, "warmstart_train.json"), "w") as f:
                json.dump(result_extra, f, indent=1)
        return result_extra
    test_params = sweep.sample_params(test_n, seed + 1)
    pred = None
    budget = step_budget(train_steps)
    if (Ln, Mn_) != (L, M):
        # A network of a coarser grid: its prediction interpolated onto this
        # grid; the step budget scales with the grid's cold step counts
        # (not known here: none).
        pred = lambda: predict_on_grid(net, norm, test_params, Ln, Mn_, L, M)  # noqa: E731
        budget = None
    result, _, _ = evaluate(net, norm, test_params, L, M, plain_sample=plain_sample, pred=pred,
                            budget=budget, warm_levels=warm_levels, ref_threshold=ref_threshold,
                            ref_max_steps=ref_max_steps, max_steps=case_max_steps, staged=staged)
    result.update(result_extra)
    result["n_params"] = net.n_params()
    print(json.dumps(result, indent=1), flush=True)
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "warmstart_result.json"), "w") as f:
            json.dump(result, f, indent=1)
    return result
