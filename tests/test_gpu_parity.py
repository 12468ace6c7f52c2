"""GPU parity: the HIP library against the CPU oracle on the same configs.

Bar: bit-identical results (NaN matches NaN) for every conserved and physical
field, t, dt and the step count.  Where the reference itself is not
bit-determined (SolutionChange and the conserved integrals sum in OpenMP /
tree order) the tolerance is written in the test.  All arithmetic is FP64;
the stated parity tolerance of the LF step is 1e-12 relative L-infinity per
field, which bit-identity meets with room to spare.
"""
import numpy as np
import pytest

import _oracle as O
import cases
from parity import describe_mismatch, rel_linf, run_oracle, same_bits

pytestmark = pytest.mark.gpu


def _compare_state(sol, o, label):
    u, p = sol.download()
    for k in range(8):
        assert same_bits(u[k], o.u0(k)), f"{label} u0_{k + 1}: {describe_mismatch(u[k], o.u0(k))}"
    for k in range(10):
        assert same_bits(p[k], o.phys(k)), f"{label} {O.PHYS_NAMES[k]}: {describe_mismatch(p[k], o.phys(k))}"


CASES = {
    "straight_40x16": lambda **kw: cases.channel_config(40, 16, max_steps=60, **kw),
    "condiv_noise_70x130": lambda **kw: cases.channel_config(70, 130, geometry="condiv", noise=1e-3,
                                                             max_steps=40, **kw),
    "thruster_80x40": lambda **kw: cases.thruster_config(80, 40, max_steps=60, **kw),
    "mixed_24x14": lambda **kw: cases.mixed_config(24, 14, max_steps=30, **kw),
    "minimal_3x2": lambda **kw: cases.channel_config(3, 2, max_steps=12, **kw),
}


@pytest.mark.parametrize("kernel,fast_div", [("simple", False), ("simple", True), ("tiled", False),
                                             ("tiled", True)])
@pytest.mark.parametrize("name", list(CASES))
def test_run_bit_exact(mpd, name, kernel, fast_div):
    text = CASES[name](kernel=kernel, fast_div=fast_div)
    print(f"case {name} kernel={kernel} fast_div={fast_div}", flush=True)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        assert res["steps"] == ores["steps"]
        assert res["t"] == ores["t"], (res["t"], ores["t"])
        assert res["dt"] == ores["dt"], (res["dt"], ores["dt"])
        _compare_state(sol, o, name)


def test_tiled_odd_tile_rows(mpd):
    """Tile heights that do not divide L, one-row tiles and a strip count
    that leaves a mostly empty last strip."""
    for h in (1, 5, 7, 64):
        text = cases.channel_config(37, 125, geometry="condiv", noise=1e-3, max_steps=9, tile_rows=h)
        print(f"tile_rows={h}", flush=True)
        o, _ = run_oracle(text, 0)
        with mpd.Solver(text) as sol:
            sol.run()
            _compare_state(sol, o, f"tile_rows={h}")


@pytest.mark.parametrize("order,rows", [(1, 1), (1, 3), (1, 8), (2, 1), (2, 2), (2, 5), (2, 8), (2, 37), (3, 1),
                                        (3, 3), (3, 8), (3, 37)])
def test_tiled_xcd_band_order(mpd, order, rows):
    """Tile order 1 (whole row bands per XCD, workgroups past the last chunk
    or strip idle) and order 2 (four vertically adjacent tiles per workgroup,
    alternate waves marching up) give the same bits, including chunk counts
    that are not a multiple of 4 or 8, strip counts that are not a multiple
    of 4, and a single chunk (three of the four waves idle); order 3 (a
    group of strips per XCD, chunk by chunk; 5 strips here, so three XCDs
    have none)."""
    text = cases.channel_config(37, 300, geometry="condiv", noise=1e-3, max_steps=9, tile_rows=rows)
    print(f"tile_order={order} tile_rows={rows}", flush=True)
    o, _ = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        sol.set_option("tile_order", order)
        sol.run()
        _compare_state(sol, o, f"tile_order={order} rows={rows}")


@pytest.mark.parametrize("order,rows", [(0, 8), (0, 3), (2, 5)])
def test_tiled_pointer_path(mpd, order, rows):
    """The tiled stencil's plain-pointer path (what states of 2^31 bytes or
    more take; forced here with the "buffers" option): stores behind the lane
    masks instead of out-of-range buffer offsets, rows 1 and L to the row
    buffer through a pointer.  Same bits as the oracle."""
    text = cases.channel_config(37, 300, geometry="condiv", noise=1e-3, max_steps=9, tile_rows=rows)
    print(f"buffers=0 tile_order={order} tile_rows={rows}", flush=True)
    o, _ = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        sol.set_option("buffers", 0)
        sol.set_option("tile_order", order)
        sol.run()
        _compare_state(sol, o, f"buffers=0 order={order} rows={rows}")


def _spike_config(L=60, M=24, steps=80):
    """Straight channel with a one-cell v_z spike at (z, r) = (0.5, 0.5): the
    LF averaging flattens it, so the max wave speed falls by far more than
    0.1 % per step for many steps."""
    text = cases.channel_config(L, M, max_steps=steps)
    spike = '"0.1 + 40 * max(0, 1 - 300 * abs(z - 0.5)) * max(0, 1 - 250 * abs(r - 0.5))"'
    assert text.count("  v_z: 0.1\n") == 1
    return text.replace("  v_z: 0.1\n", f"  v_z: {spike}\n")


def _oracle_speed_drops(text, steps):
    """Steps at which the oracle's max wave speed falls below 0.999 x the
    previous step's: each needs the exact rescan of the whole state."""
    case = cases.oracle_case(text)
    o = O.Oracle(case)
    t, dt = 0.0, case["dt"]
    prev = None
    drops = []
    for k in range(1, steps + 1):
        res = o.run(1, t=t, dt=dt)
        t, dt = res["t"], res["dt"]
        s = o.max_wave_speed()
        if prev is not None and s < 0.999 * prev:
            drops.append(k)
        prev = s
    return o, t, dt, drops


@pytest.mark.parametrize("per_graph", [0, 6])
def test_wave_speed_rescan_in_graphs(mpd, per_graph):
    """Steps whose candidate max cannot be proven idle the rest of their graph;
    the host runs the exact rescan, which finishes the step, and launches the
    remaining steps.  The result is the oracle's, bit for bit."""
    steps = 80
    text = _spike_config(steps=steps)
    o, t, dt, drops = _oracle_speed_drops(text, steps)
    assert len(drops) >= 10, drops
    print(f"rescan steps: {len(drops)} of {steps}, steps_per_graph={per_graph}", flush=True)
    with mpd.Solver(text) as sol:
        if per_graph:
            sol.set_option("steps_per_graph", per_graph)
        res = sol.run()
        assert res["steps"] == steps and res["t"] == t and res["dt"] == dt
        _compare_state(sol, o, "spike")


@pytest.mark.parametrize("per_graph", [0, 4])
def test_wave_speed_rescan_in_timed_loop(mpd, per_graph):
    """The benchmark loop (mpd_time_steps: warm-up, timed graphs, per-kernel
    pass) with rescans; afterwards the state is the oracle's after all its
    steps."""
    warm, steps = 5, 10
    text = _spike_config(steps=0)
    o, t, dt, drops = _oracle_speed_drops(text, warm + 2 * steps)
    assert len(drops) >= 10, drops
    with mpd.Solver(text) as sol:
        if per_graph:
            sol.set_option("steps_per_graph", per_graph)
        tm = sol.time_steps(steps, warm)
        assert tm["total_ms"] > 0
        _compare_state(sol, o, "spike timed")


def test_advance_and_compute_dt_match_oracle(mpd):
    """Solver::Advance(dt) + Diagnostics::ComputeDt driven from the host, as
    src/main.cpp:225-231 drives them."""
    text = cases.thruster_config(60, 30, max_steps=0)
    o = O.Oracle(cases.oracle_case(text))
    with mpd.Solver(text) as sol:
        dt = dt_o = 1.0e-7
        for step in range(25):
            sol.advance(dt)
            o.advance(dt_o)
            dt = sol.compute_dt(dt)
            dt_o = o.compute_dt(dt_o)
            assert dt == dt_o, (step, dt, dt_o)
        assert sol.max_wave_speed() == o.max_wave_speed()
        _compare_state(sol, o, "advance")


def test_run_resumes_across_calls(mpd):
    """Several run() calls (graph relaunches with eager work in between) give
    the same state as one call."""
    text = cases.channel_config(50, 20, max_steps=0, geometry="condiv")
    o = O.Oracle(cases.oracle_case(text))
    ores = o.run(77)
    with mpd.Solver(text) as sol:
        for n in (1, 10, 3, 40, 23):
            sol.run(n)
            sol.integrals()
            sol.outlet()
        u, _ = sol.download(phys=False)
        for k in range(8):
            assert same_bits(u[k], o.u0(k)), describe_mismatch(u[k], o.u0(k))


def test_T_truncation_and_stop(mpd):
    """The last step is cut to T and the loop stops at t == T (src/main.cpp:220-223)."""
    text = cases.channel_config(30, 12, max_steps=0, T=0.0123456, dt=1e-3)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        assert res["steps"] == ores["steps"] and res["t"] == ores["t"] == 0.0123456
        again = sol.run()
        assert again["steps"] == 0
        _compare_state(sol, o, "T-truncation")


def test_convergence_loop(mpd):
    """Convergence check every check_frequency steps, snapshot at init,
    stop when SolutionChange < threshold (src/main.cpp:150-152, 233-247)."""
    text = cases.channel_config(24, 10, max_steps=4000, threshold=2e-4, check_frequency=25)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        assert res["converged"] == ores["converged"]
        assert res["steps"] == ores["steps"]
        # The residual sums in a different order than the oracle: 1e-12 relative.
        assert abs(res["change"] - ores["change"]) <= 1e-12 * abs(ores["change"])
        _compare_state(sol, o, "convergence")


def test_diagnostics(mpd):
    text = cases.channel_config(64, 40, geometry="condiv", noise=1e-3, max_steps=30)
    o, _ = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        sol.run()
        assert sol.max_wave_speed() == o.max_wave_speed()
        mf, th = sol.outlet()
        assert mf == o.mass_flux() and th == o.thrust()
        assert sol.radial_symmetry() == o.radial_symmetry()
        # Conserved integrals: per-row tree sums vs the reference's serial sum.
        ig, og = sol.integrals(), o.integrals()
        assert np.all(np.abs(ig - og) <= 1e-13 * np.maximum(np.abs(og), 1e-300) + 1e-300), (ig, og)
        o.save_prev()
        g = sol.grid()
        assert same_bits(g["r"], o.r) and same_bits(g["r_z"], o.r_z) and same_bits(g["dr"], o.dr)


def test_solution_change_explicit(mpd):
    text = cases.channel_config(32, 20, max_steps=0)
    o = O.Oracle(cases.oracle_case(text))
    o.save_prev()
    with mpd.Solver(text) as sol:
        first = sol.solution_change(save_prev=True)  # takes the IC snapshot
        assert first == 0.0
        for _ in range(10):
            sol.advance(1e-4)
            o.advance(1e-4)
        c, co = sol.solution_change(save_prev=True), o.solution_change()
        assert abs(c - co) <= 1e-12 * co


def test_upload_download_roundtrip(mpd):
    text = cases.channel_config(20, 10, max_steps=5)
    with mpd.Solver(text) as sol:
        u, p = sol.download()
        u2 = [a.copy() for a in u]
        sol.upload(u2, p)
        ub, pb = sol.download()
        for k in range(8):
            assert same_bits(ub[k], u[k])
        for k in range(10):
            assert same_bits(pb[k], p[k])


@pytest.mark.slow
def test_straight_channel_2000_steps(mpd):
    """The CPU-runnable workload in full: straight channel 200x80, 2000 steps,
    all 8 fields within 1e-12 relative L-infinity of the oracle (and in fact
    bit-identical)."""
    text = cases.channel_config(200, 80, max_steps=2000)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        assert res["steps"] == 2000 and res["t"] == ores["t"]
        u, _ = sol.download(phys=False)
        for k in range(8):
            assert rel_linf(u[k], o.u0(k)) <= 1e-12
            assert same_bits(u[k], o.u0(k)), describe_mismatch(u[k], o.u0(k))


@pytest.mark.slow
def test_full_size_6400x2560(mpd):
    """The memory-bound workload at full size: two steps bit-identical to the
    oracle, and the tiled kernel identical to the simple one after 6 more."""
    text = cases.channel_config(6400, 2560, geometry="condiv", noise=1e-3, max_steps=0)
    o = O.Oracle(cases.oracle_case(text))
    o.run(2)
    with mpd.Solver(text) as sol:
        sol.run(2)
        u, _ = sol.download(phys=False)
        for k in range(8):
            assert same_bits(u[k], o.u0(k)), describe_mismatch(u[k], o.u0(k))
        del o
        sol.run(6)
        ut, _ = sol.download(phys=False)
    with mpd.Solver(text.replace("kernel: tiled", "kernel: simple")) as sol2:
        sol2.run(8)
        us, _ = sol2.download(phys=False)
    for k in range(8):
        assert same_bits(ut[k], us[k]), describe_mismatch(ut[k], us[k])
        assert np.isfinite(ut[k][1:-1, 1:-1]).all()


@pytest.mark.slow
def test_shipped_thruster_800x400(mpd):
    """The reference's shipped case (config.yaml: 800x400 short cathode,
    3-segment M_LO with the axis_lf axis, expression BCs) for 500 steps of
    its adaptive time loop: every field, t and dt bit-identical to the
    oracle."""
    text = cases.thruster_config(800, 400, max_steps=500)
    print("shipped thruster 800x400, 500 steps", flush=True)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        assert res["steps"] == ores["steps"] == 500
        assert res["t"] == ores["t"] and res["dt"] == ores["dt"]
        _compare_state(sol, o, "thruster_800x400")


def test_run_done_flag(mpd):
    """mpd_run's done reports the time loop's own end (T, convergence or
    time.max_steps), not the call's step cap."""
    text = cases.channel_config(20, 9, max_steps=30)
    with mpd.Solver(text) as sol:
        r = sol.run(10)
        assert r["steps"] == 10 and r["done"] == 0
        r = sol.run()
        assert r["steps"] == 20 and r["total_steps"] == 30 and r["done"] == 1


def _with_ic_vphi(text, value):
    old = "  v_r: 0.0\n  v_phi: 0.0\n  H_z: 0.0"
    assert text.count(old) == 1
    return text.replace(old, f"  v_r: 0.0\n  v_phi: {value}\n  H_z: 0.0")


@pytest.mark.parametrize("vphi,label", [("1.0e-40", "IC outside the range"),
                                        ("2.0e-29", "state leaves the range during the run")])
def test_split_division_range_switch(mpd, vphi, label):
    """The split division equals IEEE division only while the state stays in
    its range (|u_1| in [2^-30, 2^31), other conserved values zero or in
    [2^-100, 2^61)).  A tiny azimuthal velocity puts u_4 below 2^-100: at
    upload (1e-40), or after a few steps as the zero inflow value eats into
    a 2e-29 field.  The solver then runs with IEEE division -- the result
    reports it -- and stays bit-identical to the oracle; the same case
    batched with an ordinary one switches the batch."""
    text = _with_ic_vphi(cases.channel_config(40, 16, geometry="condiv", noise=1e-3, max_steps=60), vphi)
    o, ores = run_oracle(text, 0)
    with mpd.Solver(text) as sol:
        res = sol.run()
        u, p = sol.download()
    print(label, res, flush=True)
    assert res["fast_div"] == 0
    assert (res["steps"], res["t"], res["dt"]) == (ores["steps"], ores["t"], ores["dt"])
    for k in range(8):
        assert same_bits(u[k], o.u0(k)), f"u0_{k + 1}: {describe_mismatch(u[k], o.u0(k))}"
    plain = cases.channel_config(30, 12, max_steps=60)
    with mpd.Solver(plain) as sol:
        assert sol.run()["fast_div"] == 1
    with mpd.Batch([text, plain]) as b:
        rb = b.run()
        assert all(r["fast_div"] == 0 for r in rb)
        ub, _ = b.instance(0).download(phys=False)
    for k in range(8):
        assert same_bits(ub[k], o.u0(k)), f"batched u0_{k + 1}"
