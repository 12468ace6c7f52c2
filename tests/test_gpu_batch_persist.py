"""GPU parity of the persistent batched loop (batch option "persistent"):
one workgroup per instance runs K steps of it in one launch.  Instance by
instance the results must be the oracle's bit for bit (fields, t, dt, step
count), whatever the mix, and the same bits as the graph of per-phase
launches, SolutionChange included (the persistent kernel sums the single
solver's block partition in the same order)."""
import numpy as np
import pytest

import _oracle as O
import cases
import sweep
from parity import describe_mismatch, run_oracle, same_bits
from test_gpu_batch import MIX, _check_batch, _compare, _conv_config, _spike

pytestmark = pytest.mark.gpu


def _persist(per_graph=0, rows=0):
    def conf(b):
        b.set_option("persistent", 1)
        if per_graph:
            b.set_option("steps_per_graph", per_graph)
        if rows:
            b.set_option("tile_rows", rows)
    return conf


def test_persist_uneven_bit_exact(mpd):
    """The seven-instance mix (3x2 to 200x80, every BC kind, step caps, an
    instance that needs exact wave-speed rescans) in persistent launches."""
    _check_batch(mpd, MIX, _persist())


@pytest.mark.parametrize("per_graph,rows", [(2, 3), (10, 1), (6, 8)])
def test_persist_launch_lengths_and_tiles(mpd, per_graph, rows):
    """Short and odd launch lengths and tile heights: the same bits."""
    _check_batch(mpd, MIX[:4] + MIX[6:], _persist(per_graph, rows))


def test_persist_convergence_and_rescans(mpd):
    """Convergence checks on their own steps (one instance whose rescans
    shift its phase, checks every 7 steps; a converging one every 10), the
    rest running on: stops, states and SolutionChange against the oracle."""
    texts = [_spike(60, 24, 50, threshold=1e-30, check_frequency=7),
             _conv_config(16, 6, "1.0", "0.1", 2e-3, 10),
             _conv_config(20, 8, "0.7", "0.2", 1e-3, 20),
             cases.channel_config(40, 16, max_steps=45, threshold=1e-30, check_frequency=9)]
    with mpd.Batch(texts) as b:
        b.set_option("persistent", 1)
        res = b.run()
        for i, t in enumerate(texts):
            o, ores = run_oracle(t, 0)
            print(f"instance {i}: {res[i]['steps']} steps, change {res[i]['change']:.6e}", flush=True)
            assert res[i]["steps"] == ores["steps"], (i, res[i], ores)
            assert bool(res[i]["converged"]) == ores["converged"], (i, res[i], ores)
            assert res[i]["t"] == ores["t"] and res[i]["dt"] == ores["dt"], (i, res[i], ores)
            assert abs(res[i]["change"] - ores["change"]) <= 1e-12 * abs(ores["change"]), (i, res[i], ores)
            _compare(b.instance(i), o, f"instance {i}")


def _run_mode(mpd, texts, persistent, steps, warm=None):
    with mpd.Batch(texts) as b:
        if warm is not None:
            for i, u in enumerate(warm):
                b.warm_start(i, u)
        b.set_option("persistent", persistent)
        res = b.run(steps)
        states = [b.instance(i).download() for i in range(len(texts))]
    return res, states


def test_persist_matches_graph_path_on_sweep_cases(mpd):
    """24 seeded sweep cases at 96x39 and 48x19 (the sweep's expression
    boundary program, convergence checks every 100 steps), 700 steps and
    then to a tight threshold's check: every instance's state, t, dt, step
    count and SolutionChange equal the graph path's bit for bit."""
    params = sweep.sample_params(24, 5)
    texts = [sweep.case_config(p, 96 if i % 2 == 0 else 48, 39 if i % 2 == 0 else 19, threshold=1e-5,
                               max_steps=20000) for i, p in enumerate(params)]
    for steps in (700, 0):
        rg, sg = _run_mode(mpd, texts, 0, steps)
        rp, sp = _run_mode(mpd, texts, 1, steps)
        for i in range(len(texts)):
            for key in ("steps", "t", "dt", "change", "converged", "done"):
                assert rg[i][key] == rp[i][key], (steps, i, key, rg[i], rp[i])
            for k in range(8):
                assert same_bits(sg[i][0][k], sp[i][0][k]), f"case {i} u0_{k + 1}"
            for k in range(10):
                assert same_bits(sg[i][1][k], sp[i][1][k]), f"case {i} {O.PHYS_NAMES[k]}"
        print(f"{steps or 'to the end'}: steps {[r['steps'] for r in rp][:6]} ..., "
              f"converged {sum(r['converged'] for r in rp)}/{len(rp)}", flush=True)


def test_persist_warm_start_matches_oracle_continuation(mpd):
    """Warm starts (uploaded states) then persistent launches: each instance
    equals the graph path from the same warm state, and a cold-started
    instance in the same batch equals the oracle."""
    texts = [cases.channel_config(40, 16, max_steps=0, threshold=1e-6, check_frequency=10),
             cases.channel_config(30, 12, max_steps=0, rho="1.3", threshold=1e-6, check_frequency=10),
             cases.channel_config(24, 10, max_steps=90)]
    warm = []
    for t in texts[:2]:
        with mpd.Solver(t) as s:
            s.run(35)
            warm.append(s.download(phys=False)[0])
    rg, sg = _run_mode(mpd, texts[:2], 0, 120, warm)
    rp, sp = _run_mode(mpd, texts[:2], 1, 120, warm)
    for i in range(2):
        assert rg[i] == rp[i], (i, rg[i], rp[i])
        for k in range(8):
            assert same_bits(sg[i][0][k], sp[i][0][k]), f"warm {i} u0_{k + 1}"
    with mpd.Batch(texts) as b:
        for i, u in enumerate(warm):
            b.warm_start(i, u)
        b.set_option("persistent", 1)
        res = b.run()
        o, ores = run_oracle(texts[2], 0)
        assert (res[2]["steps"], res[2]["t"], res[2]["dt"]) == (ores["steps"], ores["t"], ores["dt"])
        _compare(b.instance(2), o, "cold instance")
