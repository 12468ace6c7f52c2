"""GPU test of mpd_create_with_grid: a wall profile computed in the caller's
process (a cubic inner wall and a cosine outer wall, in no geometry registry
of the library), passed as the per-row values Grid::Build evaluates an
IGeometry at (src/grid.cpp:20-42), against the oracle given the same rows:
grid arrays, then 200 steps of the sweep case's boundary program (slip-wall
expressions read r_z), bit for bit."""
import numpy as np
import pytest

import _oracle as O
import cases
from parity import describe_mismatch, same_bits

pytestmark = pytest.mark.gpu


def wall_rows(L):
    z = (np.arange(L + 2) - 1.0) * (1.0 / L)
    r_in = 0.22 + 0.06 * z ** 3
    r_out = 0.85 - 0.04 * np.cos(2.0 * np.pi * z)
    dr_in = 0.18 * z ** 2
    dr_out = 0.08 * np.pi * np.sin(2.0 * np.pi * z)
    return r_in, r_out, dr_in, dr_out


def test_host_grid_matches_oracle(mpd):
    import sweep
    L, M, steps = 96, 40, 200
    rows = wall_rows(L)
    text = sweep.case_config([1.3, 0.2, 0.3, 0.0, 0.0], L, M, threshold=0.0, max_steps=steps)
    case = cases.oracle_case(text)
    case["geom_rows"] = rows
    o = O.Oracle(case)
    with mpd.Solver(text, grid_rows=rows) as sol:
        g = sol.grid()
        assert same_bits(g["r"], o.r) and same_bits(g["r_z"], o.r_z)
        assert same_bits(g["R"], o.R) and same_bits(g["dr"], o.dr) and g["r_0"] == o.r0
        u_ic, p_ic = sol.download()
        for k in range(8):
            assert same_bits(u_ic[k], o.u0(k)), f"IC u0_{k + 1}"
        res = sol.run()
        ores = o.run(steps)
        assert (res["steps"], res["t"], res["dt"]) == (ores["steps"], ores["t"], ores["dt"]), (res, ores)
        u, p = sol.download()
    for k in range(8):
        assert same_bits(u[k], o.u0(k)), f"u0_{k + 1}: {describe_mismatch(u[k], o.u0(k))}"
    for k in range(10):
        assert same_bits(p[k], o.phys(k)), O.PHYS_NAMES[k]


def test_host_grid_rejects_bad_rows(mpd):
    import sweep
    L, M = 20, 8
    r_in, r_out, dr_in, dr_out = wall_rows(L)
    text = sweep.case_config([1.0, 0.1, 0.2, 0.0, 0.0], L, M)
    bad = r_out.copy()
    bad[3] = np.nan
    with pytest.raises(mpd.MpdError) as e:
        mpd.Solver(text, grid_rows=(r_in, bad, dr_in, dr_out))
    assert e.value.code == mpd.MPD_E_ARG
