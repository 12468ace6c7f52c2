"""Workload configs shared by the tests, bench.py and the smoke check.

Each case is config.yaml text in the reference's schema
(include/config.hpp:119-177).  `oracle_case(text)` turns the same text into
the oracle's parameter dict through PyYAML, a parse path independent of the
library's own YAML reader.
"""
import yaml

import _oracle as O

GAMMA_53 = "1.6666666666666667"

_SLIP_WALLS = """
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

_INFLOW_OUTFLOW = """
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
"""


def channel_config(L=200, M=80, *, geometry="straight", max_steps=2000, dt=1.0e-4, cfl=0.4, rho="1.0",
                   uz="0.1", amp="0.2", noise=0.0, seed=0, kernel="tiled", fast_div=True, T=1.0e9,
                   threshold=0.0, check_frequency=100, a_inner=0.05, a_outer=0.1, tile_rows=0):
    """The straight coaxial channel (the CPU-runnable workload: z in [0,1],
    r in [0.2,0.8], inflow fixed at z=0, zero-gradient outflow, slip walls) or
    the converging-diverging channel r_in = 0.2 + a_in sin(pi z),
    r_out = 0.8 - a_out z^2 of the memory-bound workloads."""
    if geometry == "straight":
        geom = "geometry:\n  type: straight_coaxial\n  params: { r_inner: 0.2, r_outer: 0.8 }\n"
    else:
        geom = ("geometry:\n  type: converging_diverging\n"
                f"  params: {{ r_inner: 0.2, a_inner: {a_inner}, r_outer: 0.8, a_outer: {a_outer} }}\n")
    noise_txt = f"  rho_noise: {{ amplitude: {noise}, seed: {seed} }}\n" if noise else ""
    conv = f"convergence:\n  threshold: {threshold}\n  check_frequency: {check_frequency}\n" if threshold else ""
    return f"""
physics:
  gamma: {GAMMA_53}
  beta: 0.05
  H_z0: 0.25
time:
  T: {T}
  dt: {dt}
  max_steps: {max_steps}
adaptive_dt:
  enabled: true
  cfl_number: {cfl}
  growth_factor: 1.1
  dt_min: 1.0e-12
  dt_max: 1.0e-2
{conv}grid:
  L_max: {L}
  M_max: {M}
{geom}initial_conditions:
  rho: {rho}
  v_z: {uz}
  v_r: 0.0
  v_phi: 0.0
  H_z: 0.0
  H_r: 0.0
  H_phi: "{amp} / r"
  e: "0.5 / ((gamma - 1) * rho)"
{noise_txt}{_INFLOW_OUTFLOW.format(rho=rho, uz=uz, amp=amp)}{_SLIP_WALLS}gpu:
  kernel: {kernel}
  fast_div: {"true" if fast_div else "false"}
  tile_rows: {tile_rows}
"""


def thruster_config(L=80, M=40, max_steps=200, kernel="tiled", fast_div=True):
    """The reference's shipped thruster case (config.yaml) at a reduced grid:
    cosine cathode geometry, 3-segment M_LO with cathode wall, injection
    segment and axis_lf axis, expression BCs.  Ranges scaled from 800 rows."""
    a = int(round(235 * L / 800))
    b = int(round(263 * L / 800))
    return f"""
physics:
  gamma: 1.67
  beta: 0.01
  H_z0: 0.25
time:
  T: 6
  dt: 1.0e-7
  max_steps: {max_steps}
adaptive_dt:
  enabled: true
  cfl_number: 0.45
  growth_factor: 1.10
  dt_min: 1.0e-10
  dt_max: 5.0e-5
grid:
  L_max: {L}
  M_max: {M}
geometry:
  type: short_cathode_cosine
  params:
    z_center: 0.31
    transition_half_width: 0.02
    r_inner_before: 0.200
    r_inner_after: 0.005
    r_outer: 0.800
initial_conditions:
  rho: 0.1
  v_z: 0.1
  v_r: 0.1
  v_phi: 0.0
  H_z: H_z0
  H_r: "H_z * r_z"
  H_phi: "(1 - 0.9 * z) * r_0 / r"
  e: "beta / (2 * (gamma - 1))"
boundary_conditions:
  l_lo:
    - rho: neumann
      v_z: {{ dirichlet: 0.0 }}
      v_r: neumann
      v_phi: neumann
      e: neumann
      H_z: "H_z0"
      H_r: {{ dirichlet: 0.0 }}
      H_phi: "r_0 / r"
  l_hi:
    - rho: neumann
  m_hi:
    - v_r: "v_z * r_z"
      H_r: "H_z * r_z"
  m_lo:
    - range: [0, {a}]
      v_r: "v_z_nb * r_z_nb"
      H_r: "H_z_nb * r_z_nb"
    - range: [{a + 1}, {b}]
      rho: {{ dirichlet: 0.1 }}
      v_phi: {{ dirichlet: 0.0 }}
      e: "beta / (2 * (gamma - 1))"
      H_r: "H_z_nb * r_z_nb"
    - range: [{b + 1}, -1]
      rho: axis_lf
      v_z: axis_lf
      v_r: {{ dirichlet: 0.0 }}
      v_phi: {{ dirichlet: 0.0 }}
      e: axis_lf
      H_z: axis_lf
      H_r: {{ dirichlet: 0.0 }}
      H_phi: {{ dirichlet: 0.0 }}
gpu:
  kernel: {kernel}
  fast_div: {"true" if fast_div else "false"}
"""


def mixed_config(L=24, M=14, max_steps=40, kernel="tiled", fast_div=True):
    """Every BC feature at once: Neumann offsets, Dirichlet, expressions over
    neighbour/geometry symbols, overlapping segments, a partly uncovered face
    (its cells go NaN as in the reference), AxisLF next to a Neumann v_z."""
    return f"""
physics: {{ gamma: 1.4, beta: 0.1, H_z0: 0.3 }}
time: {{ T: 1.0e9, dt: 2.0e-4, max_steps: {max_steps} }}
adaptive_dt: {{ enabled: true, cfl_number: 0.3, growth_factor: 1.05, dt_min: 1.0e-12, dt_max: 1.0e-2 }}
grid: {{ L_max: {L}, M_max: {M} }}
geometry:
  type: converging_diverging
  params: {{ r_inner: 0.25, a_inner: 0.04, r_outer: 0.9, a_outer: 0.12 }}
initial_conditions:
  vars: {{ amp: 0.7, k: 3.0 }}
  rho: "1 + 0.2 * sin(k * pi * z) * (r - r_0)"
  v_z: "0.05 + 0.02 * cos(pi * r)"
  v_r: "0.01 * v_z"
  v_phi: "0.03 * z"
  H_z: "H_z0 * (1 + z)"
  H_r: "H_z * r_z"
  H_phi: "amp * r_0 / r"
  e: "beta / (2 * (gamma - 1)) + 0.1 * exp(-z)"
  rho_noise: {{ amplitude: 0.01, seed: 7 }}
boundary_conditions:
  l_lo:
    - range: [0, 5]
      rho: {{ dirichlet: 1.1 }}
      v_z: {{ neumann: 0.01 }}
      H_phi: "r_0 / r"
      e: "beta / (2 * (gamma - 1) * rho)"
    - range: [4, -1]
      rho: {{ type: neumann, gradient: -0.001 }}
      v_z: {{ type: dirichlet, value: 0.06 }}
      v_r: {{ expr: "v_r_nb * 0.5 + dz * dr" }}
      H_phi: "H_phi_nb * r_nb / r"
  l_hi:
    - rho: neumann
      v_z: "max(v_z_nb, 0.0)"
      e: "sqrt(e_nb * e_nb)"
      H_z: "H_z_nb + (z - 1) * 0"
  m_hi:
    - range: [0, {L - 4}]
      v_r: "v_z * r_z"
      H_r: "H_z * r_z"
      v_phi: dirichlet
  m_lo:
    - range: [0, {L // 2}]
      v_r: "v_z_nb * r_z_nb"
      H_r: "H_z_nb * r_z_nb"
      rho: {{ neumann: 0.002 }}
    - range: [{L // 2 - 2}, -1]
      rho: axis_lf
      v_z: neumann
      e: axis_lf
      H_z: axis_lf
      v_r: {{ dirichlet: 0.0 }}
      H_r: {{ dirichlet: 0.0 }}
gpu:
  kernel: {kernel}
  fast_div: {"true" if fast_div else "false"}
"""


# --------------------------------------------------------------- oracle side

_GEOM = {"straight_coaxial": (O.GEOM_STRAIGHT, ["r_inner", "r_outer"], [0.2, 0.8]),
         "converging_diverging": (O.GEOM_CONDIV, ["r_inner", "a_inner", "r_outer", "a_outer"],
                                  [0.2, 0.05, 0.8, 0.1]),
         "short_cathode_cosine": (O.GEOM_CATHODE_COSINE,
                                  ["z_center", "transition_half_width", "r_inner_before", "r_inner_after",
                                   "r_outer"], [0.31, 0.015, 0.2, 0.005, 0.8])}


def _num(v):
    return repr(float(v)) if isinstance(v, (int, float)) and not isinstance(v, bool) else str(v)


def _cond(v):
    """The field-condition grammar of src/config.cpp:52-202."""
    if v is None or v == "neumann":
        return ("neumann", 0.0)
    if v == "dirichlet":
        return ("dirichlet", 0.0)
    if v == "axis_lf":
        return ("axis_lf",)
    if isinstance(v, str):
        return ("expr", v)
    if isinstance(v, (int, float)):
        return ("expr", _num(v))
    if isinstance(v, dict):
        if "neumann" in v:
            return ("neumann", float(v["neumann"]))
        if "dirichlet" in v:
            return ("dirichlet", float(v["dirichlet"]))
        if "expr" in v and "type" not in v:
            return ("expr", str(v["expr"]))
        t = v["type"]
        if t == "neumann":
            return ("neumann", float(v.get("gradient", v.get("value", 0.0))))
        if t == "dirichlet":
            return ("dirichlet", float(v.get("value", 0.0)))
        if t == "axis_lf":
            return ("axis_lf",)
        if t in ("expression", "expr"):
            return ("expr", str(v["expr"]))
    raise ValueError(v)


def oracle_case(text):
    d = yaml.safe_load(text)
    ph = d.get("physics", {})
    tm = d.get("time", {})
    ad = d.get("adaptive_dt", {})
    cv = d.get("convergence", {})
    gr = d["grid"]
    g = d.get("geometry", {})
    gt, keys, defaults = _GEOM[g.get("type", "short_cathode_cosine")]
    params = g.get("params") or {}
    gp = [float(params.get(k, dv)) for k, dv in zip(keys, defaults)]
    ic = dict(d.get("initial_conditions") or {})
    noise = ic.pop("rho_noise", None) or {}
    ic_vars = ic.pop("vars", None) or {}
    case = dict(L=int(gr["L_max"]), M=int(gr["M_max"]),
                gamma=float(ph.get("gamma", 1.67)), beta=float(ph.get("beta", 0.05)),
                H_z0=float(ph.get("H_z0", 0.25)), T=float(tm.get("T", 0.5)), dt=float(tm.get("dt", 2.5e-5)),
                adaptive=bool(ad.get("enabled", False)), cfl=float(ad.get("cfl_number", 0.5)),
                growth=float(ad.get("growth_factor", 1.1)), dt_min=float(ad.get("dt_min", 1e-9)),
                dt_max=float(ad.get("dt_max", 1e-3)), conv_threshold=float(cv.get("threshold", 0.0)),
                check_frequency=int(cv.get("check_frequency", 100)), geom=(gt, gp),
                ic={k: _num(v) for k, v in ic.items() if v is not None},
                ic_vars={k: float(v) for k, v in ic_vars.items()},
                noise_amp=float(noise.get("amplitude", 0.0)), noise_seed=int(noise.get("seed", 0)),
                max_steps=int(tm.get("max_steps", 0)))
    faces = {}
    for f in ("l_lo", "l_hi", "m_lo", "m_hi"):
        segs = []
        for item in (d.get("boundary_conditions") or {}).get(f) or []:
            sg = {"range": tuple(item.get("range", (-1, -1)))}
            for fld in O.BC_FIELDS:
                sg[fld] = _cond(item.get(fld))
            segs.append(sg)
        faces[f] = segs
    case["faces"] = faces
    return case
