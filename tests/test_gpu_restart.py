"""GPU runs from a VTK restart (VtkIC, src/ics/vtk_ic.cpp:42-239) against
the oracle from the same initial state.

A GPU run writes a frame (IOManager::WriteVtk, src/io_manager.cpp:218-315);
a restart config reads it back, at the same size (a continuation) or mapped
onto another grid.  The restart's physical arrays (the host VtkIC, whose
mapping tests/test_vtk.py pins on the CPU) are handed to the oracle, whose
own InitConservative must give the same conserved state as the library's;
then the GPU loop from the restart config and the oracle loop run the same
steps and must agree bit for bit (fields, t, dt, step count)."""
import numpy as np
import pytest

import _oracle as O
import cases
from parity import describe_mismatch, same_bits

pytestmark = pytest.mark.gpu


def _with_vtk_ic(text, path, exact):
    i = text.index("initial_conditions:")
    j = text.index("boundary_conditions:")
    return text[:i] + f"initial_conditions:\n  vtk_file: \"{path}\"\n  vtk_exact_z: {str(exact).lower()}\n" + text[j:]


@pytest.mark.parametrize("geometry,L0,M0,L1,M1,exact", [
    ("straight", 40, 16, 40, 16, True),            # continuation at the same size
    ("straight", 40, 16, 57, 23, False),           # onto a finer grid, the reference's axial mapping
    ("converging_diverging", 48, 20, 30, 12, True),  # onto a coarser grid
])
def test_restart_run_matches_oracle(mpd, tmp_path, geometry, L0, M0, L1, M1, exact):
    print(f"restart {geometry} {L0}x{M0} -> {L1}x{M1} exact={exact}", flush=True)
    first = cases.channel_config(L0, M0, geometry=geometry, noise=1e-3, seed=4)
    with mpd.Solver(first) as s:
        s.run(150)
        _, ph = s.download(u0=False)
    path = tmp_path / "frame.vtk"
    mpd.write_vtk_fields(first, ph, path)
    base = cases.channel_config(L1, M1, geometry=geometry)
    cfg = _with_vtk_ic(base, path, exact)
    u_ic, ph_ic = mpd.initial_state(cfg)
    assert all(np.all(np.isfinite(a)) for a in ph_ic) and np.all(ph_ic[O.RHO][1:L1 + 1, 1:M1 + 2] > 0)

    o = O.Oracle(cases.oracle_case(base))
    for k in range(10):
        o.phys(k)[...] = ph_ic[k]
    o.init_conservative()
    for k in range(8):
        assert same_bits(o.u0(k), u_ic[k]), f"InitConservative u0_{k + 1}: {describe_mismatch(o.u0(k), u_ic[k])}"
    steps = 300
    ores = o.run(steps)
    with mpd.Solver(cfg) as s:
        res = s.run(steps)
        u, p = s.download()
    assert (res["steps"], res["t"], res["dt"]) == (ores["steps"], ores["t"], ores["dt"]), (res, ores)
    for k in range(8):
        assert same_bits(u[k], o.u0(k)), f"u0_{k + 1}: {describe_mismatch(u[k], o.u0(k))}"
    for k in range(10):
        assert same_bits(p[k], o.phys(k)), f"{O.PHYS_NAMES[k]}: {describe_mismatch(p[k], o.phys(k))}"
