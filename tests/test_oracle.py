"""Pins of the CPU oracle (oracle/lf_oracle.c), CPU only.

1. Golden fixtures written by the reference's own compiled solver code
   (tests/golden/make_golden.py): bit-identical.
2. The reference harness itself, when built here (oracle/_ref): bit-identical
   on further cases.
3. Closed forms and invariants the mathematics fixes.
4. A pure-Python restatement of a full step, boundary conditions included,
   on tiny grids (brute force): bit-identical.  This is what pins the
   boundary-condition half of the oracle (src/bcs/per_field_bc.cpp needs
   exprtk and cannot be built here).
"""
import glob
import math
import os

import numpy as np
import pytest

import _oracle as O
import cases
import pyref
from parity import same_bits

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("digest_"))


def _oracle_from_state(L, M, geom, gp, gamma, phys, u0, **dt_params):
    o = O.Oracle(dict(L=L, M=M, gamma=gamma, geom=(geom, list(gp)), **dt_params))
    for k in range(10):
        o.phys(k)[:] = phys[k]
    for k in range(8):
        o.u0(k)[:] = u0[k]
    return o


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_fixture(path):
    z = np.load(path)
    L, M = int(z["param_L"]), int(z["param_M"])
    gp = list(z["param_gp"])
    o = _oracle_from_state(L, M, int(z["param_geom"]), gp, float(z["param_gamma"]), z["in_phys"], z["in_u0"],
                           cfl=float(z["param_cfl"]), growth=float(z["param_growth"]),
                           dt_min=float(z["param_dt_min"]), dt_max=float(z["param_dt_max"]))
    assert same_bits(o.r, z["r"]) and same_bits(o.r_z, z["r_z"])
    assert same_bits(o.dr, z["dr"]) and same_bits(o.R, z["R"]) and o.r0 == float(z["r0"])
    o.save_prev()
    assert o.max_wave_speed() == float(z["speed0"])
    assert same_bits(o.integrals(), z["integrals0"])
    dt = float(z["param_dt"])
    dts = []
    for _ in range(int(z["param_steps"])):
        o.advance(dt)
        dts.append(dt)
        if int(z["param_adaptive"]):
            dt = o.compute_dt(dt)
    assert same_bits(np.array(dts), z["dts"]) and dt == float(z["dt_next"])
    for k in range(8):
        assert same_bits(o.u0(k), z["out_u0"][k]), f"u0_{k + 1}"
    for k in range(10):
        assert same_bits(o.phys(k), z["out_phys"][k]), O.PHYS_NAMES[k]
    assert same_bits(o.integrals(), z["integrals"])
    assert same_bits(o.mass_flux(), z["mass_flux"]) and same_bits(o.thrust(), z["thrust"])
    sp = o.max_wave_speed()
    assert sp == float(z["speed"]) or (math.isnan(sp) and math.isnan(float(z["speed"])))
    # SolutionChange: the reference sums in OpenMP order, the oracle by rows.
    c, cz = o.solution_change(), float(z["change"])
    assert (math.isnan(c) and math.isnan(cz)) or abs(c - cz) <= 1e-13 * abs(cz)


def _ref_case(L, M, seed, steps, adaptive):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    phys, u0 = mg.random_state(L, M, seed)
    gp = [0.22, 0.03, 0.85, 0.07, 0.0]
    out = O.run_ref_harness(L, M, O.GEOM_CONDIV, gp, 1.55, steps, 3e-4, adaptive, 0.35, 1.07, 1e-12, 1e-2,
                            phys, u0)
    o = _oracle_from_state(L, M, O.GEOM_CONDIV, gp[:4], 1.55, phys, u0, cfl=0.35, growth=1.07, dt_min=1e-12,
                           dt_max=1e-2)
    dt = 3e-4
    for _ in range(steps):
        o.advance(dt)
        if adaptive:
            dt = o.compute_dt(dt)
    return o, out, dt


@pytest.mark.skipif(not O.ref_harness_available(), reason="reference harness not built (no reference tree)")
@pytest.mark.parametrize("L,M,seed,steps,adaptive", [(16, 11, 1, 1, 0), (7, 30, 2, 4, 1), (25, 9, 3, 2, 1)])
def test_oracle_matches_reference_harness(L, M, seed, steps, adaptive):
    o, out, dt = _ref_case(L, M, seed, steps, adaptive)
    assert dt == out["dt_next"]
    for k in range(8):
        assert same_bits(o.u0(k), out["u0"][k]), f"u0_{k + 1}"
    for k in range(10):
        assert same_bits(o.phys(k), out["phys"][k]), O.PHYS_NAMES[k]


# ------------------------------------------------------------ closed forms


def _uniform_case(L=12, M=9, vz=0.0, hphi=0.0, **kw):
    ic = {"rho": "1.3", "v_z": repr(vz), "v_r": "0.0", "v_phi": "0.0", "H_z": "0.0", "H_r": "0.0",
          "H_phi": repr(hphi), "e": "0.7"}
    allneu = [{f: ("neumann", 0.0) for f in O.BC_FIELDS}]
    c = dict(L=L, M=M, gamma=1.4, geom=(O.GEOM_STRAIGHT, [0.3, 0.9]), ic=ic,
             faces={"l_lo": allneu, "l_hi": allneu, "m_lo": allneu, "m_hi": allneu})
    c.update(kw)
    return c


def test_uniform_quiescent_state_is_steady():
    """A uniform state at rest with H = 0 in a straight channel with
    zero-gradient walls is a steady solution: the u_3 source P cancels the
    pressure flux -d(rP)/dr exactly in exact arithmetic, so only rounding
    remains after many steps."""
    o = O.Oracle(_uniform_case())
    before = [o.phys(k)[1:-1, 1:-1].copy() for k in range(10)]
    for _ in range(50):
        o.advance(1e-3)
    for k in (O.RHO, O.E, O.P, O.PT):
        a = o.phys(k)[1:-1, 1:-1]
        assert np.max(np.abs(a - before[k])) <= 1e-12 * np.max(np.abs(before[k]))
    for k in (O.VZ, O.VR, O.VPHI, O.HZ, O.HR, O.HPHI):
        assert np.max(np.abs(o.phys(k)[1:-1, 1:-1])) <= 1e-12


def test_outlet_mass_flux_closed_form():
    """For uniform rho*v_z the trapezoid rule is exact for the linear integrand
    2 pi r: mass flux = rho v_z pi (r_out^2 - r_in^2)."""
    o = O.Oracle(_uniform_case(vz=0.25, M=17))
    exact = 1.3 * 0.25 * math.pi * (0.9 ** 2 - 0.3 ** 2)
    assert abs(o.mass_flux() - exact) <= 1e-14 * exact
    p = (1.4 - 1.0) * 1.3 * 0.7
    exact_t = (1.3 * 0.25 * 0.25 + p) * math.pi * (0.9 ** 2 - 0.3 ** 2)
    assert abs(o.thrust() - exact_t) <= 1e-14 * exact_t


def test_integrals_closed_form():
    """sum over cells of rho r dr dz = rho dr Mn (r_in + r_out) / 2 for a
    uniform density (r is linear in m)."""
    L, M = 12, 9
    o = O.Oracle(_uniform_case(L, M))
    dr = (0.9 - 0.3) / M
    mass = 1.3 * dr * (M + 1) * (0.3 + 0.9) / 2.0
    q = o.integrals()
    assert abs(q[0] - mass) <= 1e-13 * mass
    assert abs(q[4] - 0.7 * mass) <= 1e-13 * mass  # energy = rho e r ...
    assert q[2] == 0.0 and q[3] == 0.0 and q[5] == 0.0 and q[6] == 0.0


def test_wave_speed_and_dt_closed_form():
    o = O.Oracle(_uniform_case(vz=0.2, hphi=0.3, cfl=0.5, growth=1.1, dt_min=1e-9, dt_max=1e-3))
    p = (1.4 - 1.0) * 1.3 * 0.7
    s = math.sqrt(0.2 * 0.2) + math.sqrt(1.4 * p / 1.3) + math.sqrt(0.3 * 0.3 / 1.3)
    assert abs(o.max_wave_speed() - s) <= 2e-16 * s
    dx = min(1 / 12, 1 / 9)
    assert o.compute_dt(1.0) == min(max(0.5 * dx / (o.max_wave_speed() + 1e-10), 1e-9), 1e-3)
    assert o.compute_dt(1e-6) == 1e-6 * 1.1  # growth cap
    o2 = O.Oracle(_uniform_case(vz=0.2, hphi=0.3, cfl=0.5, growth=1.1, dt_min=1e-9, dt_max=1e-7))
    assert o2.compute_dt(1.0) == 1e-7  # clamp


def test_conservative_physical_roundtrip():
    """u = f^-1(phys) then phys' = f(u) reproduces phys within two ulps."""
    o = O.Oracle(cases.oracle_case(cases.mixed_config()))
    ph = [o.phys(k).copy() for k in range(10)]
    o.u(0)[:] = o.u0(0)
    for k in range(8):
        o.u(k)[:] = o.u0(k)
    o.update_physical(1, o.rows - 2, 1, o.cols - 2)
    for k in (O.RHO, O.VZ, O.VR, O.VPHI, O.E, O.HZ, O.HR, O.HPHI):
        a, b = o.phys(k)[1:-1, 1:-1], ph[k][1:-1, 1:-1]
        assert np.all(np.abs(a - b) <= 4.5e-16 * np.abs(b) + 1e-300)


def test_neumann_and_dirichlet_boundary_values():
    """After a step the physical values of boundary cells are f(rebuilt u):
    the Neumann neighbour value (+offset) / the Dirichlet value to an ulp."""
    faces = {"l_lo": [dict({f: ("dirichlet", 0.5 + 0.1 * i) for i, f in enumerate(O.BC_FIELDS)})],
             "l_hi": [dict({f: ("neumann", 0.01) for f in O.BC_FIELDS})],
             "m_lo": [dict({f: ("neumann", 0.0) for f in O.BC_FIELDS})],
             "m_hi": [dict({f: ("neumann", 0.0) for f in O.BC_FIELDS})]}
    c = _uniform_case(10, 8, vz=0.1, hphi=0.2)
    c["faces"] = faces
    o = O.Oracle(c)
    o.advance(1e-4)
    L = 10
    ids = {"rho": O.RHO, "v_z": O.VZ, "v_r": O.VR, "v_phi": O.VPHI, "H_z": O.HZ, "H_r": O.HR, "H_phi": O.HPHI,
           "e": O.E}
    for i, f in enumerate(O.BC_FIELDS):
        row = o.phys(ids[f])[1, 1:-1]
        v = 0.5 + 0.1 * i
        assert np.all(np.abs(row - v) <= 3e-16 * v), f
    # L_HI: neighbour row L-1 (its final value) + 0.01, compared to 2 ulps.
    for f in ("rho", "e"):
        a = o.phys(ids[f])[L, 2:-2]
        b = o.phys(ids[f])[L - 1, 2:-2] + 0.01
        assert np.all(np.abs(a - b) <= 5e-16 * np.abs(b)), f


# -------------------------------------------------------------- brute force

BRUTE = {
    "straight": lambda: cases.channel_config(6, 4, max_steps=3),
    "condiv_noise": lambda: cases.channel_config(7, 5, geometry="condiv", noise=1e-2, max_steps=3),
    "thruster": lambda: cases.thruster_config(12, 6, max_steps=3),
    "mixed": lambda: cases.mixed_config(9, 6, max_steps=3),
}


@pytest.mark.parametrize("name", list(BRUTE))
def test_oracle_matches_python_restatement(name):
    text = BRUTE[name]()
    case = cases.oracle_case(text)
    o = O.Oracle(case)
    ref = pyref.PyRef(case, [o.phys(k).copy() for k in range(10)], [o.u0(k).copy() for k in range(8)],
                      o.r.copy(), o.r_z.copy(), o.dr.copy(), o.r0)
    dt = case["dt"]
    dt_o = dt
    for _ in range(3):
        ref.advance(dt)
        o.advance(dt_o)
        dt = ref.compute_dt(dt)
        dt_o = o.compute_dt(dt_o)
        assert dt == dt_o
    for k in range(8):
        assert same_bits(o.u0(k), ref.u0[k]), f"u0_{k + 1}"
    for k in range(10):
        assert same_bits(o.phys(k), ref.ph[k]), O.PHYS_NAMES[k]


def test_noise_generator_known_values():
    """SplitMix64 (Steele, Lea, Flood 2014) of seed 0: the first outputs of the
    published generator with state 0 are 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4;
    U = (z >> 11) 2^-53."""
    lib = O.lib()
    assert lib.orc_noise_u01(0, 0) == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert lib.orc_noise_u01(0, 1) == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53


def test_oracle_grid_from_rows():
    """Per-row wall data (the caller's IGeometry evaluated at Grid::Build's
    points): constant rows give the straight channel's grid bit for bit, and
    a cubic profile gives r = (1 - frac) r_in + frac r_out, r_z likewise,
    R = r_out - r_in, dr = R / M_max and r_0 from row 1 (z = 0), in
    src/grid.cpp:20-42's operation order."""
    L, M = 12, 7
    straight = O.Oracle(dict(L=L, M=M, geom=(O.GEOM_STRAIGHT, [0.3, 0.9])))
    const = [np.full(L + 2, 0.3), np.full(L + 2, 0.9), np.zeros(L + 2), np.zeros(L + 2)]
    rows = O.Oracle(dict(L=L, M=M, geom_rows=const))
    for a, b in ((rows.r, straight.r), (rows.r_z, straight.r_z), (rows.R, straight.R), (rows.dr, straight.dr)):
        assert same_bits(a, b)
    assert rows.r0 == straight.r0
    z = (np.arange(L + 2) - 1.0) * (1.0 / L)
    rin, rout = 0.2 + 0.05 * z ** 3, 0.8 - 0.03 * z
    drin, drout = 0.15 * z ** 2, np.full(L + 2, -0.03)
    o = O.Oracle(dict(L=L, M=M, geom_rows=[rin, rout, drin, drout]))
    frac = (np.arange(M + 3) - 1.0) * (1.0 / M)
    assert same_bits(o.r, (1.0 - frac)[None, :] * rin[:, None] + frac[None, :] * rout[:, None])
    assert same_bits(o.r_z, (1.0 - frac)[None, :] * drin[:, None] + frac[None, :] * drout[:, None])
    assert same_bits(o.R, rout - rin) and same_bits(o.dr, (rout - rin) / M)
    assert o.r0 == (rin[1] + rout[1]) / 2.0
