"""Writes the golden fixtures of tests/golden/ from the reference's own solver
code (oracle/_ref/ref_harness, built by `make -C oracle ref` from the
reference sources in this container).

The reference ships no tests, fixtures or known answers for the LF step, so
these vectors are the reference pins of the oracle.  The harness links the
reference's solver.cpp, fields.cpp, grid.cpp, face_bc.cpp, diagnostics.cpp,
integrals.cpp and verification.cpp; boundary segments are left empty (the
BC translation unit needs exprtk, which is not available), so the boundary
columns of these runs turn to NaN exactly as the reference's do without BCs.

Each fixture stores the input state (10 physical + 8 conserved arrays), the
run parameters and every harness output.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as O  # noqa: E402


def random_state(L, M, seed):
    """Independent random physical and conserved arrays (interior only; ghost
    ring zero): every stencil term sees distinct values."""
    rng = np.random.default_rng(seed)
    rows, cols = L + 2, M + 3
    phys = [np.zeros((rows, cols)) for _ in range(10)]
    u0 = [np.zeros((rows, cols)) for _ in range(8)]
    inner = (slice(1, rows - 1), slice(1, cols - 1))
    lo_hi = {O.RHO: (0.5, 2.0), O.E: (0.2, 1.0), O.P: (0.1, 1.0), O.PT: (0.2, 1.5)}
    for k in range(10):
        a, b = lo_hi.get(k, (-0.5, 0.5))
        phys[k][inner] = rng.uniform(a, b, size=(L, M + 1))
    for k in range(8):
        a, b = (0.3, 1.5) if k == 0 else (-0.5, 0.5)
        u0[k][inner] = rng.uniform(a, b, size=(L, M + 1))
    return phys, u0


def ic_state(case):
    o = O.Oracle(case)
    return [o.phys(k).copy() for k in range(10)], [o.u0(k).copy() for k in range(8)]


FIXTURES = {
    # One step from a random state on the converging-diverging grid.
    "random_condiv_10x7_1step": dict(
        L=10, M=7, geom=O.GEOM_CONDIV, gp=[0.2, 0.05, 0.8, 0.1], gamma=1.6666666666666667, steps=1,
        dt=1.0e-3, adaptive=0, cfl=0.4, growth=1.1, dt_min=1e-12, dt_max=1e-2, state="random", seed=11),
    # Three adaptive steps from another random state.
    "random_condiv_9x12_3steps": dict(
        L=9, M=12, geom=O.GEOM_CONDIV, gp=[0.25, 0.04, 0.9, 0.12], gamma=1.4, steps=3,
        dt=2.0e-4, adaptive=1, cfl=0.3, growth=1.05, dt_min=1e-12, dt_max=1e-2, state="random", seed=5),
    # The straight-channel IC, 20 adaptive steps (boundary columns go NaN
    # without BCs; the NaN front moves one cell per step).
    "straight_ic_14x40_20steps": dict(
        L=14, M=40, geom=O.GEOM_STRAIGHT, gp=[0.2, 0.8], gamma=1.6666666666666667, steps=20,
        dt=1.0e-4, adaptive=1, cfl=0.4, growth=1.1, dt_min=1e-12, dt_max=1e-2, state="ic"),
}

IC = {"rho": "1.0", "v_z": "0.1", "v_r": "0.0", "v_phi": "0.0", "H_z": "0.0", "H_r": "0.0",
      "H_phi": "0.2 / r", "e": "0.5 / ((gamma - 1) * rho)"}


def make(name, f):
    if f["state"] == "random":
        phys, u0 = random_state(f["L"], f["M"], f["seed"])
    else:
        case = dict(L=f["L"], M=f["M"], gamma=f["gamma"], geom=(f["geom"], f["gp"]), ic=IC)
        phys, u0 = ic_state(case)
    out = O.run_ref_harness(f["L"], f["M"], f["geom"], f["gp"], f["gamma"], f["steps"], f["dt"], f["adaptive"],
                            f["cfl"], f["growth"], f["dt_min"], f["dt_max"], phys, u0)
    arrays = dict(in_phys=np.stack(phys), in_u0=np.stack(u0), out_u0=np.stack(out["u0"]),
                  out_phys=np.stack(out["phys"]), r=out["r"], r_z=out["r_z"], R=out["R"], dr=out["dr"],
                  dts=out["dts"])
    scalars = {k: np.float64(out[k]) for k in ("r0", "speed0", "dt_next", "change", "speed", "mass_flux",
                                               "thrust")}
    params = {k: np.asarray(v) for k, v in f.items() if k != "state"}
    np.savez_compressed(os.path.join(HERE, name + ".npz"), integrals0=out["integrals0"], integrals=out["integrals"],
                        **arrays, **scalars, **{"param_" + k: v for k, v in params.items()})
    print("wrote", name)


if __name__ == "__main__":
    if not O.ref_harness_available():
        sys.exit("oracle/_ref/ref_harness is missing: run `make -C oracle ref` in the build container")
    for n, f in FIXTURES.items():
        make(n, f)
