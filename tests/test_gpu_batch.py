"""GPU parity of the batched sweep: instances of uneven size advanced together
by the batched kernels must give, instance by instance, the oracle's results
bit for bit (every field, t, dt, step count), whatever the batch mixes
(grid sizes, BC programs, step caps, convergence stops, wave-speed rescans),
the tile heights the autotuner picks and the launch grid."""
import numpy as np
import pytest

import _oracle as O
import cases
from parity import describe_mismatch, run_oracle, same_bits

pytestmark = pytest.mark.gpu


def _compare(sol, o, label):
    u, p = sol.download()
    for k in range(8):
        assert same_bits(u[k], o.u0(k)), f"{label} u0_{k + 1}: {describe_mismatch(u[k], o.u0(k))}"
    for k in range(10):
        assert same_bits(p[k], o.phys(k)), f"{label} {O.PHYS_NAMES[k]}: {describe_mismatch(p[k], o.phys(k))}"


def _spike(L, M, steps, **kw):
    text = cases.channel_config(L, M, max_steps=steps, **kw)
    spike = '"0.1 + 40 * max(0, 1 - 300 * abs(z - 0.5)) * max(0, 1 - 250 * abs(r - 0.5))"'
    return text.replace("  v_z: 0.1\n", f"  v_z: {spike}\n")


MIX = [
    ("straight_40x16", lambda: cases.channel_config(40, 16, max_steps=60)),
    ("condiv_noise_70x130", lambda: cases.channel_config(70, 130, geometry="condiv", noise=1e-3, max_steps=40)),
    ("thruster_80x40", lambda: cases.thruster_config(80, 40, max_steps=60)),
    ("mixed_24x14", lambda: cases.mixed_config(24, 14, max_steps=30)),
    ("minimal_3x2", lambda: cases.channel_config(3, 2, max_steps=12)),
    ("inflow_200x80", lambda: cases.channel_config(200, 80, max_steps=25, rho="1.7", uz="0.25", amp="0.35")),
    ("spike_60x24", lambda: _spike(60, 24, 50)),
]


def _check_batch(mpd, named, configure=None, max_steps=0):
    texts = [f() for _, f in named]
    with mpd.Batch(texts) as b:
        if configure:
            configure(b)
        res = b.run(max_steps)
        for i, (name, _) in enumerate(named):
            print(f"instance {i} {name}: {res[i]['steps']} steps", flush=True)
            o, ores = run_oracle(texts[i], max_steps if max_steps else 0)
            assert res[i]["steps"] == ores["steps"], (name, res[i], ores)
            assert res[i]["t"] == ores["t"] and res[i]["dt"] == ores["dt"], (name, res[i], ores)
            _compare(b.instance(i), o, name)
    return res


def test_batch_uneven_bit_exact(mpd):
    """Seven instances: grid sizes from 3x2 to 200x80, every BC kind, step
    caps that end the instances at different graph steps, and one instance
    whose wave speed falls fast enough to need exact rescans."""
    _check_batch(mpd, MIX)


@pytest.mark.parametrize("grid,per_graph,rows", [(7, 6, 0), (0, 2, 3), (1, 10, 1)])
def test_batch_grid_and_graph_shapes(mpd, grid, per_graph, rows):
    """A persistent stencil grid far smaller than the work list, short graphs
    and odd tile heights give the same bits."""
    def conf(b):
        b.set_option("grid", grid)
        b.set_option("steps_per_graph", per_graph)
        if rows:
            b.set_option("tile_rows", rows)
    _check_batch(mpd, MIX[:4] + MIX[6:], conf)


def test_batch_autotune_then_run(mpd):
    """The autotuner's picks (per grid size tile heights, stencil grid) run
    bit-identically; its report names every grid size."""
    def conf(b):
        rep = b.autotune(reps=3)
        print(rep, flush=True)
        assert "tile_rows" in rep and "stencil grid" in rep
        for i in range(b.n):
            assert b.tile_rows(i) in (2, 3, 4, 6, 8, 12, 16, 24, 32)
    _check_batch(mpd, MIX, conf)


def test_batch_max_steps_cap_and_resume(mpd):
    """mpd_batch_run's own step cap, then a second call that finishes each
    instance's loop: the same bits as one uninterrupted oracle run."""
    named = MIX[:3]
    texts = [f() for _, f in named]
    with mpd.Batch(texts) as b:
        r1 = b.run(17)
        assert all(r["steps"] == 17 and not r["done"] for r in r1), r1
        r2 = b.run(0)
        for i, (name, _) in enumerate(named):
            o, ores = run_oracle(texts[i], 0)
            assert r1[i]["steps"] + r2[i]["steps"] == ores["steps"], name
            assert r2[i]["done"] and r2[i]["t"] == ores["t"] and r2[i]["dt"] == ores["dt"], name
            _compare(b.instance(i), o, name)


def _conv_config(L, M, rho, uz, thr, cf):
    return cases.channel_config(L, M, max_steps=3000, rho=rho, uz=uz, threshold=thr, check_frequency=cf)


def test_batch_convergence_stops(mpd):
    """Instances with convergence checks stop at their own check steps (the
    reference's loop, src/main.cpp:233-247), leave the work lists, and the
    rest run on; states, step counts and SolutionChange match the oracle."""
    specs = [(16, 6, "1.0", "0.1", 2e-3, 10), (20, 8, "0.7", "0.2", 1e-3, 20), (12, 5, "1.5", "0.3", 5e-3, 10),
             (24, 10, "1.0", "0.05", 0.0, 100)]
    texts = [_conv_config(*s) for s in specs]
    texts[3] = cases.channel_config(24, 10, max_steps=400, uz="0.05")
    oracles = []
    for t in texts:
        case = cases.oracle_case(t)
        o = O.Oracle(case)
        res = o.run(case["max_steps"])
        oracles.append((o, res))
    stops = [r["steps"] for _, r in oracles]
    print("oracle steps:", stops, flush=True)
    assert len(set(stops)) >= 3, stops
    with mpd.Batch(texts) as b:
        res = b.run()
        for i, (o, ores) in enumerate(oracles):
            assert res[i]["steps"] == ores["steps"], (i, res[i], ores)
            assert bool(res[i]["converged"]) == ores["converged"], (i, res[i], ores)
            assert res[i]["t"] == ores["t"] and res[i]["dt"] == ores["dt"]
            if ores["converged"]:
                # SolutionChange sums in a different order than the oracle.
                assert abs(res[i]["change"] - ores["change"]) <= 1e-12 * abs(ores["change"]), (res[i], ores)
            _compare(b.instance(i), o, f"instance {i}")


def test_batch_warm_start_matches_single_solver(mpd):
    """A warm start (upload of u0 with physical = f(u0), snapshot retaken)
    then a batched run equals the single-instance solver doing the same."""
    a = cases.channel_config(40, 16, max_steps=0, threshold=1e-6, check_frequency=10)
    c = cases.channel_config(30, 12, max_steps=0, rho="1.3", threshold=1e-6, check_frequency=10)
    warm = []
    for t in (a, c):
        with mpd.Solver(t) as s:
            s.run(35)
            u, _ = s.download(phys=False)
            warm.append(u)
    ref = []
    for t, u in zip((a, c), warm):
        with mpd.Solver(t) as s:
            s.upload(u)
            s.solution_change(save_prev=True)
            r = s.run(120)
            ref.append((r, s.download()))
    with mpd.Batch([a, c]) as b:
        for i, u in enumerate(warm):
            b.warm_start(i, u)
        res = b.run(120)
        for i, (r, (ru, rp)) in enumerate(ref):
            assert res[i]["steps"] == r["steps"] and res[i]["t"] == r["t"] and res[i]["dt"] == r["dt"]
            # SolutionChange: the batched kernels sum over the single solver's
            # block partition in its order, so the value is the same bits.
            assert res[i]["change"] == r["change"], (res[i], r)
            u, p = b.instance(i).download()
            for k in range(8):
                assert same_bits(u[k], ru[k]), f"u0_{k + 1}: {describe_mismatch(u[k], ru[k])}"
            for k in range(10):
                assert same_bits(p[k], rp[k]), f"phys {k}: {describe_mismatch(p[k], rp[k])}"


def test_batch_instances_are_borrowed(mpd):
    """Instance views stay valid while the batch lives; destroying a view
    leaves the instance in place (mpd_destroy ignores batch instances)."""
    texts = [cases.channel_config(20, 8, max_steps=5), cases.channel_config(24, 10, max_steps=7)]
    with mpd.Batch(texts) as b:
        v = b.instance(1)
        v.close()
        res = b.run()
        assert [r["steps"] for r in res] == [5, 7]
        assert b.instance(1).download()[0][0].shape == (26, 13)


def test_batch_checks_across_rescans(mpd):
    """Graphs carry the convergence launch only on steps where some instance
    checks.  An instance whose step waits for an exact wave-speed rescan
    (finished on the host) changes phase against the graph; its later checks
    (every 7 steps, threshold never met) and a converging instance's (every
    10 steps) must still run on the oracle's steps: same SolutionChange after
    the last check, same stop, same bits."""
    texts = [_spike(60, 24, 50, threshold=1e-30, check_frequency=7),
             _conv_config(16, 6, "1.0", "0.1", 2e-3, 10),
             cases.channel_config(40, 16, max_steps=45, threshold=1e-30, check_frequency=9)]
    with mpd.Batch(texts) as b:
        res = b.run()
        for i, t in enumerate(texts):
            o, ores = run_oracle(t, 0)
            print(f"instance {i}: {res[i]['steps']} steps, change {res[i]['change']:.6e}", flush=True)
            assert res[i]["steps"] == ores["steps"], (i, res[i], ores)
            assert bool(res[i]["converged"]) == ores["converged"], (i, res[i], ores)
            assert res[i]["t"] == ores["t"] and res[i]["dt"] == ores["dt"], (i, res[i], ores)
            # SolutionChange sums in a different order than the oracle, in
            # the single solver's order (the same bits as a run alone).
            assert abs(res[i]["change"] - ores["change"]) <= 1e-12 * abs(ores["change"]), (i, res[i], ores)
            with mpd.Solver(t) as s:
                solo = s.run()
            assert (res[i]["steps"], res[i]["change"]) == (solo["steps"], solo["change"]), (i, res[i], solo)
            _compare(b.instance(i), o, f"instance {i}")
