"""Pins of the oracle's boundary-condition half that the mathematics fixes
(CPU only; test infrastructure).

The reference's boundary code (src/bcs/per_field_bc.cpp) needs exprtk and
cannot be compiled here, so these checks pin the oracle's restatement of it
without a second restatement:

1. Face order and corner ownership (src/solver.cpp:29-32: L_LO, M_HI, M_LO,
   L_HI; src/face_bc.cpp:56-67 face ranges; :118-124 the M_LO corner policy).
   Each face gets values no other face can produce (distinct Dirichlet
   values, distinct Neumann offsets), so the owner of every ring cell and the
   state each Neumann read sees can be read off the final physical fields.
2. The AxisLF half-stencils (src/bcs/per_field_bc.cpp:490-566) on fields that
   are polynomial in l and linear in m, against closed forms derived from
   the discrete operators: the central z difference of a product of two
   linear functions is exact, the one-sided radial difference over dr_l of a
   linear function is its slope.  Each test also shows that the plausible
   misreadings (2 dr_l instead of dr_l, U1's radial velocity at m instead of
   m + 1, the neighbour's density without the 1/r weight, ...) land far
   outside the tolerance.
3. AxisLF inside a full step: which state each read sees (u0 of the start
   of the step, the central update's v_r at m = 2, the face value of v_r at
   m = 1, the start-of-step p), on a state whose z terms vanish exactly.
"""
import math

import numpy as np
import pytest

import _oracle as O

EPS = 2.0 ** -52
BC_ID = {"rho": O.RHO, "v_z": O.VZ, "v_r": O.VR, "v_phi": O.VPHI, "H_z": O.HZ, "H_r": O.HR, "H_phi": O.HPHI,
         "e": O.E}

IC = {"rho": "1 + 0.2 * z + 0.3 * r", "v_z": "0.1 + 0.05 * z", "v_r": "0.01 * r", "v_phi": "0.02",
      "H_z": "0.03 * r", "H_r": "0.01", "H_phi": "0.2 / r", "e": "0.5 / ((gamma - 1) * rho)"}


def _case(faces, L=9, M=7, gamma=1.4):
    return dict(L=L, M=M, gamma=gamma, geom=(O.GEOM_CONDIV, [0.25, 0.04, 0.85, 0.08]), ic=dict(IC), faces=faces)


def _seg(fn):
    return [{f: fn(i, f) for i, f in enumerate(O.BC_FIELDS)}]


def _close(a, b, ulps=8):
    return abs(a - b) <= ulps * EPS * max(abs(a), abs(b), 1e-300)


def _check(o, want, label):
    """want: {(l, m): {field: value}}; each final physical value within a
    few ulps (a face value goes through RebuildUFromPhysical and the final
    UpdatePhysicalFromU)."""
    for (l, m), vals in want.items():
        for f, v in vals.items():
            got = o.phys(BC_ID[f])[l, m]
            assert _close(got, v), f"{label}: {f} at ({l},{m}) = {got!r}, want {v!r}"


def test_corner_owners_distinct_dirichlet():
    """Every face Dirichlet with its own values.  Final owners: row 1 and
    (1,1) -> L_LO; (1,Mn) -> M_HI (it runs after L_LO and covers l = 1);
    column 1 for l = 2..L-1 -> M_LO (which skips l = 1 and l = L); column Mn
    -> M_HI; row L, corners included -> L_HI (last)."""
    val = {face: (lambda i, f, face=face: ("dirichlet", 1.0 + 0.1 * face + 0.01 * i))
           for face in range(4)}
    faces = {"l_lo": _seg(val[0]), "m_hi": _seg(val[1]), "m_lo": _seg(val[2]), "l_hi": _seg(val[3])}
    o = O.Oracle(_case(faces))
    L, Mn = o.rows - 2, o.cols - 2
    o.advance(1e-4)

    def owner(l, m):
        if l == L:
            return 3
        if m == Mn:
            return 1
        if l == 1:
            return 0
        return 2

    want = {}
    for l in range(1, L + 1):
        for m in range(1, Mn + 1):
            if l in (1, L) or m in (1, Mn):
                want[(l, m)] = {f: 1.0 + 0.1 * owner(l, m) + 0.01 * i for i, f in enumerate(O.BC_FIELDS)}
    assert len(want) == 2 * L + 2 * Mn - 4
    _check(o, want, "dirichlet owners")
    # A misreading that let M_LO own (1,1) or M_HI skip (1,Mn) would differ by 0.1.
    assert not _close(o.phys(O.RHO)[1, 1], 1.2) and not _close(o.phys(O.RHO)[1, Mn], 1.0)


def test_neumann_chain_through_the_faces():
    """L_LO Dirichlet, the other faces Neumann with distinct offsets.  Each
    Neumann read sees the state left by the faces before it:
    (1,Mn) = L_LO's (1,Mn-1) + o_MHI; column 1 = central (l,2) + o_MLO but
    (1,1) untouched by M_LO; row L = (L-1, m) + o_LHI, so its corners carry
    both offsets of their chain."""
    o_mhi, o_mlo, o_lhi = 0.003, 0.005, 0.007
    faces = {"l_lo": _seg(lambda i, f: ("dirichlet", 0.9 + 0.01 * i)),
             "m_hi": _seg(lambda i, f: ("neumann", o_mhi)),
             "m_lo": _seg(lambda i, f: ("neumann", o_mlo)),
             "l_hi": _seg(lambda i, f: ("neumann", o_lhi))}
    o = O.Oracle(_case(faces))
    L, Mn = o.rows - 2, o.cols - 2
    o.advance(1e-4)
    ph = {f: o.phys(BC_ID[f]).copy() for f in O.BC_FIELDS}
    want = {}
    for i, f in enumerate(O.BC_FIELDS):
        c1 = 0.9 + 0.01 * i
        for m in range(1, Mn):
            want.setdefault((1, m), {})[f] = c1
        want.setdefault((1, Mn), {})[f] = c1 + o_mhi
        for l in range(2, L):
            want.setdefault((l, Mn), {})[f] = ph[f][l, Mn - 1] + o_mhi
            want.setdefault((l, 1), {})[f] = ph[f][l, 2] + o_mlo
        for m in range(2, Mn):
            want.setdefault((L, m), {})[f] = ph[f][L - 1, m] + o_lhi
        want.setdefault((L, 1), {})[f] = (ph[f][L - 1, 2] + o_mlo) + o_lhi
        want.setdefault((L, Mn), {})[f] = (ph[f][L - 1, Mn - 1] + o_mhi) + o_lhi
    _check(o, want, "neumann chain")
    # Order pins: M_HI before L_LO would give (1,Mn) = central (1,Mn-1) + o_MHI;
    # M_LO on l = 1 would add o_MLO at (1,1); L_HI before M_LO would miss o_MLO at (L,1).
    assert not _close(o.phys(O.RHO)[1, 1], 0.9 + o_mlo)
    assert not _close(o.phys(O.RHO)[L, 1], ph["rho"][L - 1, 2] + o_lhi)


def test_neumann_reads_start_of_step_state_before_its_face_runs():
    """L faces Neumann, M faces Dirichlet.  L_LO runs first: at (1,1) it reads
    (2,1), which M_LO has not set yet, so it sees the start-of-step physical
    value (the IC on step 1); M_LO then skips l = 1.  (1,Mn) is M_HI's
    (after L_LO).  L_HI runs last and reads column values the M faces set."""
    a, b = 0.0125, -0.0075
    c_mhi, c_mlo = 0.6, 0.4
    faces = {"l_lo": _seg(lambda i, f: ("neumann", a)),
             "m_hi": _seg(lambda i, f: ("dirichlet", c_mhi + 0.01 * i)),
             "m_lo": _seg(lambda i, f: ("dirichlet", c_mlo + 0.01 * i)),
             "l_hi": _seg(lambda i, f: ("neumann", b))}
    o = O.Oracle(_case(faces))
    L, Mn = o.rows - 2, o.cols - 2
    start = {f: o.phys(BC_ID[f]).copy() for f in O.BC_FIELDS}
    o.advance(1e-4)
    ph = {f: o.phys(BC_ID[f]).copy() for f in O.BC_FIELDS}
    want = {}
    for i, f in enumerate(O.BC_FIELDS):
        want.setdefault((1, 1), {})[f] = start[f][2, 1] + a
        for m in range(2, Mn):
            want.setdefault((1, m), {})[f] = ph[f][2, m] + a
            want.setdefault((L, m), {})[f] = ph[f][L - 1, m] + b
        want.setdefault((1, Mn), {})[f] = c_mhi + 0.01 * i
        for l in range(2, L):
            want.setdefault((l, 1), {})[f] = c_mlo + 0.01 * i
            want.setdefault((l, Mn), {})[f] = c_mhi + 0.01 * i
        want.setdefault((L, 1), {})[f] = (c_mlo + 0.01 * i) + b
        want.setdefault((L, Mn), {})[f] = (c_mhi + 0.01 * i) + b
    _check(o, want, "start-of-step reads")
    assert not _close(o.phys(O.RHO)[1, 1], c_mlo + a)


def test_segment_ranges_and_config_order():
    """Two overlapping L_LO segments: the later one wins on the overlap
    (config order); negative bounds mean the face's first / last index;
    M_HI segment bounds are global l indices 0..L_max-1 (local l = g + 1)."""
    seg_a = dict({f: ("dirichlet", 0.7) for f in O.BC_FIELDS}, range=(-1, 4))
    seg_b = dict({f: ("dirichlet", 0.8) for f in O.BC_FIELDS}, range=(3, -1))
    neu = {f: ("neumann", 0.0) for f in O.BC_FIELDS}
    mhi_a = dict({f: ("dirichlet", 0.3) for f in O.BC_FIELDS}, range=(0, 2))
    mhi_b = dict({f: ("dirichlet", 0.35) for f in O.BC_FIELDS}, range=(3, -1))
    faces = {"l_lo": [seg_a, seg_b], "m_hi": [mhi_a, mhi_b], "m_lo": [dict(neu)], "l_hi": [dict(neu)]}
    o = O.Oracle(_case(faces))
    L, Mn = o.rows - 2, o.cols - 2
    o.advance(1e-4)
    rho = o.phys(O.RHO)
    for m in range(1, Mn):
        want = 0.7 if m <= 3 else 0.8  # global m = local m - 1; seg_b starts at global 3
        assert _close(rho[1, m], want), (m, rho[1, m])
    for l in range(1, L):
        want = 0.3 if l <= 3 else 0.35  # global l = local l - 1
        assert _close(rho[l, Mn], want), (l, rho[l, Mn])


# ---------------------------------------------------------------- AxisLF

# Polynomial fields: q_k = u0_k / r = A + B l + C (m - 1); v_z = V0 + V1 l;
# v_r = W0 + W1 (m - 1); H_z = Z0 + Z1 l + Z2 (m - 1); H_r = R0 + R1 (m - 1);
# P = P0 + P1 l + P2 (m - 1); p = p0 + p1 l.
COEF = dict(A=1.1, B=0.07, C=0.09, V0=0.2, V1=0.03, W0=0.05, W1=0.04, Z0=0.3, Z1=0.02, Z2=0.06, R0=0.11,
            R1=0.05, P0=0.8, P1=0.04, P2=0.03, p0=0.6, p1=0.02)
QCOEF = {0: (1.1, 0.07, 0.09), 1: (0.4, -0.03, 0.06), 4: (2.0, 0.05, -0.08), 6: (0.25, 0.02, 0.05)}


def _axis_state(L=10, M=8):
    o = O.Oracle(dict(L=L, M=M, gamma=1.4, geom=(O.GEOM_STRAIGHT, [0.3, 0.9])))
    c = COEF
    rows, cols = o.rows, o.cols
    lv = np.arange(rows, dtype=np.float64)[:, None] * np.ones((1, cols))
    mv = np.ones((rows, 1)) * (np.arange(cols, dtype=np.float64)[None, :] - 1.0)
    r = o.r
    for k, (A, B, C) in QCOEF.items():
        o.u0(k)[:] = r * (A + B * lv + C * mv)
    o.phys(O.VZ)[:] = c["V0"] + c["V1"] * lv
    o.phys(O.VR)[:] = c["W0"] + c["W1"] * mv
    o.phys(O.HZ)[:] = c["Z0"] + c["Z1"] * lv + c["Z2"] * mv
    o.phys(O.HR)[:] = c["R0"] + c["R1"] * mv
    o.phys(O.PT)[:] = c["P0"] + c["P1"] * lv + c["P2"] * mv
    o.phys(O.P)[:] = c["p0"] + c["p1"] * lv
    return o


def _closed_forms(o, l, dt):
    """AxisLF closed forms at (l, 1) and the values of plausible misreadings."""
    c = COEF
    dz = 1.0 / (o.rows - 2)
    dr = o.dr[l]
    r = o.r[l, 1]
    vz_l = c["V0"] + c["V1"] * l
    W2 = c["W0"] + c["W1"]          # v_r at m = 2
    out, wrong = {}, {}
    for k, (A, B, C) in QCOEF.items():
        qc = A + B * l               # q at (l, 1)
        qn = qc + C                  # q at (l, 2)
        avg = qc + C / 4.0           # 0.25 (q(l+1,1) + q(l-1,1) + q(l,2) + q(l,1))
        # (q(l+1,1) vz(l+1) - q(l-1,1) vz(l-1)) / (2 dz): the central
        # difference of the quadratic (A + B x)(V0 + V1 x) is its derivative.
        zadv = (B * c["V0"] + A * c["V1"] + 2.0 * B * c["V1"] * l) / dz
        radv = (qn * W2 - qc * c["W0"]) / dr   # one-sided, v_r at its own cell
        if k == 0:
            rq = C * W2 / dr                     # U1: v_r(l, m+1) in both terms
            out[k] = (avg + dt * (-zadv - rq)) * r
            wrong[k] = [(avg + dt * (-zadv - radv)) * r,            # no quirk
                        (avg + dt * (-zadv - rq / 2.0)) * r,        # 2 dr_l
                        (avg + dt * (-zadv / 2.0 - rq)) * r]        # dz instead of 2 dz
        elif k == 1:
            Hz_l = c["Z0"] + c["Z1"] * l
            # ((Hz^2 - P)(l+1) - (Hz^2 - P)(l-1)) / (2 dz) at m = 1
            zH = (2.0 * c["Z1"] * Hz_l - c["P1"]) / dz
            rH = ((Hz_l + c["Z2"]) * (c["R0"] + c["R1"]) - Hz_l * c["R0"]) / dr
            out[k] = (avg + dt * (zH + rH - zadv - radv)) * r
            wrong[k] = [(avg + dt * (zH + rH - zadv - C * W2 / dr)) * r,   # U1's quirk
                        (avg + dt * (zH + rH / 2.0 - zadv - radv / 2.0)) * r,
                        (avg + dt * (2.0 * c["Z1"] * Hz_l / dz + rH - zadv - radv)) * r]  # P dropped
        elif k == 4:
            p_l = c["p0"] + c["p1"] * l
            src = -p_l * (c["V1"] / dz + c["W1"] / dr)
            out[k] = (avg + dt * (src - zadv - radv)) * r
            wrong[k] = [(avg + dt * (-p_l * (c["V1"] / dz + c["W1"] / (2.0 * dr)) - zadv - radv)) * r,
                        (avg + dt * (-(c["P0"] + c["P1"] * l) * (c["V1"] / dz + c["W1"] / dr) - zadv - radv)) * r,
                        (avg + dt * (-zadv - radv)) * r]
        else:
            # (Hr(l,2) vz(l,2) - Hr(l,1) vz(l,1)) / dr_l, vz independent of m
            rH = c["R1"] * vz_l / dr
            out[k] = (avg + dt * (rH - radv)) * r
            wrong[k] = [(avg + dt * (rH - radv - zadv)) * r,   # a z advection term
                        (avg + dt * (rH / 2.0 - radv / 2.0)) * r]
    return out, wrong


@pytest.mark.parametrize("l", [2, 5, 8])
def test_axis_lf_closed_forms(l):
    o = _axis_state()
    dt = 0.01
    want, wrong = _closed_forms(o, l, dt)
    for k in (0, 1, 4, 6):
        got = o.axis_lf(k, l, 1, dt)
        assert abs(got - want[k]) <= 1e-13 * abs(want[k]), (k, got, want[k])
        for w in wrong[k]:
            assert abs(w - want[k]) > 1e-5 * abs(want[k]), (k, "misreading indistinguishable", w)


def test_axis_lf_in_a_step():
    """M_LO: rho, v_z, e, H_z AxisLF, the other fields Dirichlet 0; other
    faces zero-gradient.  The state is uniform in z (straight channel, IC
    independent of z), so the z terms cancel exactly and the results are
    closed forms in the values each read must see: u0 of the start of the
    step, v_r at (l,2) from the central update (the final v_r there), v_r
    at (l,1) = its Dirichlet 0, and p at (l,1) of the start of the step (the
    boundary conditions do not write p)."""
    ic = {"rho": "1 + 2 * r", "v_z": "0.15", "v_r": "0.0", "v_phi": "0.0", "H_z": "0.2 + 0.5 * r",
          "H_r": "0.0", "H_phi": "0.1 / r", "e": "1.5 - r"}
    axis = {"rho": ("axis_lf",), "v_z": ("axis_lf",), "e": ("axis_lf",), "H_z": ("axis_lf",),
            "v_r": ("dirichlet", 0.0), "v_phi": ("dirichlet", 0.0), "H_r": ("dirichlet", 0.0),
            "H_phi": ("dirichlet", 0.0)}
    neu = {f: ("neumann", 0.0) for f in O.BC_FIELDS}
    case = dict(L=8, M=6, gamma=1.4, geom=(O.GEOM_STRAIGHT, [0.1, 0.9]), ic=ic,
                faces={"l_lo": [dict(neu)], "l_hi": [dict(neu)], "m_lo": [axis], "m_hi": [dict(neu)]})
    o = O.Oracle(case)
    L = o.rows - 2
    u0 = [o.u0(k).copy() for k in range(8)]
    p_start = o.phys(O.P).copy()
    r = o.r.copy()
    dt = 2e-3
    o.advance(dt)
    vr_new = o.phys(O.VR)
    vz_new = o.phys(O.VZ)
    hr_new = o.phys(O.HR)
    for l in range(2, L):
        dr = o.dr[l]
        rc, rn = r[l, 1], r[l, 2]
        q = {k: (u0[k][l, 1] / rc, u0[k][l, 2] / rn) for k in (0, 1, 4, 6)}
        w2 = vr_new[l, 2]
        assert abs(w2) > 1e-6  # the central update moved v_r at m = 2
        want = {}
        qc, qn = q[0]
        want[0] = (0.25 * (3 * qc + qn) - dt * (qn - qc) * w2 / dr) * rc  # U1: v_r(l,2) twice
        qc, qn = q[1]
        hz_c, hz_n = 0.2 + 0.5 * rc, o.phys(O.HZ)[l, 2]
        want[1] = (0.25 * (3 * qc + qn) + dt * ((hz_n * hr_new[l, 2] - hz_c * 0.0) / dr
                                                 - (qn * w2 - qc * 0.0) / dr)) * rc
        qc, qn = q[4]
        want[4] = (0.25 * (3 * qc + qn) + dt * (-p_start[l, 1] * (w2 / dr) - (qn * w2) / dr)) * rc
        qc, qn = q[6]
        want[6] = (0.25 * (3 * qc + qn) + dt * ((hr_new[l, 2] * vz_new[l, 2]) / dr - (qn * w2) / dr)) * rc
        for k in (0, 1, 4, 6):
            got = o.u0(k)[l, 1]
            assert abs(got - want[k]) <= 1e-12 * abs(want[k]), (l, k, got, want[k])
        # Reading v_r(l,1) from the start of the step or v_r(l,2) of the
        # start of the step would give U1 = average * r:
        assert abs(o.u0(0)[l, 1] - 0.25 * (3 * q[0][0] + q[0][1]) * rc) > 1e-9 * abs(want[0])


def test_axis_lf_rejects_other_k():
    o = _axis_state()
    assert math.isnan(o.axis_lf(2, 3, 1, 0.01))
