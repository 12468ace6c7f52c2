"""A pure-Python restatement of one reference time step for tiny grids, used
to brute-force check the C oracle (test infrastructure only).

Written cell by cell from the reference: Solver::Advance
(src/solver.cpp:22-159), Fields::UpdatePhysicalFromU / CopyUToU0
(src/fields.cpp:76-116), FaceBC::Apply (src/face_bc.cpp:104-129),
PerFieldBC::Apply and the AxisLF half-stencils (src/bcs/per_field_bc.cpp:
252-566), RebuildUFromPhysical (include/physics_utils.hpp:15-24) and
Diagnostics::MaxWaveSpeed / ComputeDt (src/diagnostics.cpp:14-65).
Python floats are IEEE binary64 and every expression keeps the reference's
operation order, so results must match the oracle bit for bit.

Boundary expressions are evaluated with Python's own parser: '^' becomes
'**' (Python's power is right associative and binds tighter than unary
minus, the convention the library documents).
"""
import math

RHO, VZ, VR, VPHI, E, P, PT, HZ, HR, HPHI = range(10)
COND_FIELDS = [("rho", RHO), ("v_z", VZ), ("v_r", VR), ("v_phi", VPHI), ("H_z", HZ), ("H_r", HR),
               ("H_phi", HPHI), ("e", E)]
NB_SYMS = [("rho_nb", RHO), ("v_z_nb", VZ), ("v_r_nb", VR), ("v_phi_nb", VPHI), ("e_nb", E), ("H_z_nb", HZ),
           ("H_r_nb", HR), ("H_phi_nb", HPHI)]


def _q(a, b):
    """IEEE quotient (Python raises on a zero divisor)."""
    if b == 0.0:
        if a != a or a == 0.0:
            return math.nan
        neg = (a < 0) != (math.copysign(1.0, b) < 0)
        return -math.inf if neg else math.inf
    return a / b


def _sqrt(x):
    return math.sqrt(x) if x >= 0 else math.nan


def _sign(x):
    return 1.0 if x > 0 else (-1.0 if x < 0 else 0.0)


_FUNCS = dict(sin=math.sin, cos=math.cos, tan=math.tan, asin=math.asin, acos=math.acos, atan=math.atan,
              atan2=math.atan2, exp=math.exp, log=math.log, log2=math.log2, log10=math.log10, sqrt=math.sqrt,
              pow=math.pow, abs=math.fabs, sign=_sign, pi=math.pi,
              min=lambda *a: min(a), max=lambda *a: max(a),
              clamp=lambda lo, x, hi: lo if x < lo else (hi if x > hi else x))


class _Seg:
    def __init__(self, sg):
        self.lo, self.hi = sg["range"]
        self.conds = [sg[name] for name, _ in COND_FIELDS]
        self.code = [compile(c[1].replace("^", "**"), "<bc>", "eval") if c[0] == "expr" else None
                     for c in self.conds]
        self.has_expr = any(c[0] == "expr" for c in self.conds)
        self.sym = {}  # persists across cells, calls and steps, as exprtk's symbol values do


class PyRef:
    def __init__(self, case, phys, u0, r, r_z, dr, r0):
        self.L = case["L"]
        self.Mmax = case["M"]
        self.Mn = self.Mmax + 1
        self.rows, self.cols = self.L + 2, self.Mn + 2
        self.gamma, self.beta, self.H_z0 = case["gamma"], case["beta"], case["H_z0"]
        self.cfl, self.growth, self.dt_min, self.dt_max = case["cfl"], case["growth"], case["dt_min"], case["dt_max"]
        self.dz = 1.0 / self.L
        self.dy = 1.0 / self.Mmax
        self.r0 = r0
        self.ph = [a.astype(float).tolist() for a in phys]
        self.u0 = [a.astype(float).tolist() for a in u0]
        self.u = [[[0.0] * self.cols for _ in range(self.rows)] for _ in range(8)]
        self.r = r.tolist()
        self.rz = r_z.tolist()
        self.dr = dr.tolist()
        self.faces = [[_Seg(sg) for sg in case["faces"].get(f, [])] for f in ("l_lo", "l_hi", "m_lo", "m_hi")]

    # Fields::UpdatePhysicalFromU
    def physical(self, l_lo, l_hi, m_lo, m_hi):
        u, ph, r, g = self.u, self.ph, self.r, self.gamma
        for l in range(l_lo, l_hi + 1):
            for m in range(m_lo, m_hi + 1):
                ph[RHO][l][m] = u[0][l][m] / r[l][m]
                ph[VZ][l][m] = _q(u[1][l][m], u[0][l][m])
                ph[VR][l][m] = _q(u[2][l][m], u[0][l][m])
                ph[VPHI][l][m] = _q(u[3][l][m], u[0][l][m])
                ph[HPHI][l][m] = u[5][l][m]
                ph[HZ][l][m] = u[6][l][m] / r[l][m]
                ph[HR][l][m] = u[7][l][m] / r[l][m]
                ph[E][l][m] = _q(u[4][l][m], u[0][l][m])
                ph[P][l][m] = (g - 1.0) * ph[RHO][l][m] * ph[E][l][m]
                ph[PT][l][m] = ph[P][l][m] + 0.5 * (ph[HZ][l][m] * ph[HZ][l][m] + ph[HR][l][m] * ph[HR][l][m]
                                                    + ph[HPHI][l][m] * ph[HPHI][l][m])

    # Solver::ComputeCentralUpdate
    def central(self, dt):
        u0, u, ph, r, dz = self.u0, self.u, self.ph, self.r, self.dz
        rho, vz, vr, vphi, p, PP = ph[RHO], ph[VZ], ph[VR], ph[VPHI], ph[P], ph[PT]
        Hz, Hr, Hp = ph[HZ], ph[HR], ph[HPHI]
        for l in range(1, self.L + 1):
            for m in range(2, self.Mn):
                d = self.dr[l]
                a = [u0[k] for k in range(8)]

                def avg(k):
                    return 0.25 * (a[k][l + 1][m] + a[k][l - 1][m] + a[k][l][m + 1] + a[k][l][m - 1])

                def fz(k):
                    return a[k][l + 1][m] * vz[l + 1][m] - a[k][l - 1][m] * vz[l - 1][m]

                def fr(k):
                    return a[k][l][m + 1] * vr[l][m + 1] - a[k][l][m - 1] * vr[l][m - 1]

                u[0][l][m] = avg(0) + dt * (-fz(0) / (2 * dz) - fr(0) / (2 * d))
                u[1][l][m] = avg(1) + dt * (
                    ((Hz[l + 1][m] * Hz[l + 1][m] - PP[l + 1][m]) * r[l + 1][m]
                     - (Hz[l - 1][m] * Hz[l - 1][m] - PP[l - 1][m]) * r[l - 1][m]) / (2 * dz)
                    + (Hz[l][m + 1] * Hr[l][m + 1] * r[l][m + 1] - Hz[l][m - 1] * Hr[l][m - 1] * r[l][m - 1]) / (2 * d)
                    - fz(1) / (2 * dz) - fr(1) / (2 * d))
                u[2][l][m] = avg(2) + dt * (
                    (rho[l][m] * vphi[l][m] * vphi[l][m] + PP[l][m] - Hp[l][m] * Hp[l][m])
                    + (Hz[l + 1][m] * Hr[l + 1][m] * r[l + 1][m] - Hz[l - 1][m] * Hr[l - 1][m] * r[l - 1][m]) / (2 * dz)
                    + ((Hr[l][m + 1] * Hr[l][m + 1] - PP[l][m + 1]) * r[l][m + 1]
                       - (Hr[l][m - 1] * Hr[l][m - 1] - PP[l][m - 1]) * r[l][m - 1]) / (2 * d)
                    - fz(2) / (2 * dz) - fr(2) / (2 * d))
                u[3][l][m] = avg(3) + dt * (
                    (-rho[l][m] * vr[l][m] * vphi[l][m] + Hp[l][m] * Hr[l][m])
                    + (Hp[l + 1][m] * Hz[l + 1][m] * r[l + 1][m] - Hp[l - 1][m] * Hz[l - 1][m] * r[l - 1][m]) / (2 * dz)
                    + (Hp[l][m + 1] * Hr[l][m + 1] * r[l][m + 1] - Hp[l][m - 1] * Hr[l][m - 1] * r[l][m - 1]) / (2 * d)
                    - fz(3) / (2 * dz) - fr(3) / (2 * d))
                u[4][l][m] = avg(4) + dt * (
                    -p[l][m] * ((vz[l + 1][m] * r[l + 1][m] - vz[l - 1][m] * r[l - 1][m]) / (2 * dz)
                                + (vr[l][m + 1] * r[l][m + 1] - vr[l][m - 1] * r[l][m - 1]) / (2 * d))
                    - fz(4) / (2 * dz) - fr(4) / (2 * d))
                u[5][l][m] = avg(5) + dt * (
                    (Hz[l + 1][m] * vphi[l + 1][m] - Hz[l - 1][m] * vphi[l - 1][m]) / (2 * dz)
                    + (Hr[l][m + 1] * vphi[l][m + 1] - Hr[l][m - 1] * vphi[l][m - 1]) / (2 * d)
                    - fz(5) / (2 * dz) - fr(5) / (2 * d))
                u[6][l][m] = avg(6) + dt * (
                    (Hr[l][m + 1] * vz[l][m + 1] * r[l][m + 1] - Hr[l][m - 1] * vz[l][m - 1] * r[l][m - 1]) / (2 * d)
                    - fr(6) / (2 * d))
                u[7][l][m] = avg(7) + dt * (
                    (Hz[l + 1][m] * vr[l + 1][m] * r[l + 1][m] - Hz[l - 1][m] * vr[l - 1][m] * r[l - 1][m]) / (2 * dz)
                    - fz(7) / (2 * dz))

    def rebuild(self, l, m):
        ph, u, r = self.ph, self.u, self.r[l][m]
        u[0][l][m] = ph[RHO][l][m] * r
        u[1][l][m] = ph[RHO][l][m] * ph[VZ][l][m] * r
        u[2][l][m] = ph[RHO][l][m] * ph[VR][l][m] * r
        u[3][l][m] = ph[RHO][l][m] * ph[VPHI][l][m] * r
        u[4][l][m] = ph[RHO][l][m] * ph[E][l][m] * r
        u[5][l][m] = ph[HPHI][l][m]
        u[6][l][m] = ph[HZ][l][m] * r
        u[7][l][m] = ph[HR][l][m] * r

    def _axis_common(self, k, l, m):
        u0, r = self.u0[k], self.r
        n, s_, e, c = u0[l + 1][m] / r[l + 1][m], u0[l - 1][m] / r[l - 1][m], u0[l][m + 1] / r[l][m + 1], \
            u0[l][m] / r[l][m]
        return n, s_, e, c, 0.25 * (n + s_ + e + c)

    def axis(self, which, l, m, dt):
        vz, vr, Hz, Hr, PP, p = (self.ph[VZ], self.ph[VR], self.ph[HZ], self.ph[HR], self.ph[PT], self.ph[P])
        r, d, dz = self.r, self.dr[l], self.dz
        if which == 0:
            n, s_, e, c, a = self._axis_common(0, l, m)
            return (a + dt * (-(n * vz[l + 1][m] - s_ * vz[l - 1][m]) / (2.0 * dz)
                              - (e * vr[l][m + 1] - c * vr[l][m + 1]) / d)) * r[l][m]
        if which == 1:
            n, s_, e, c, a = self._axis_common(1, l, m)
            return (a + dt * (((Hz[l + 1][m] * Hz[l + 1][m] - PP[l + 1][m])
                               - (Hz[l - 1][m] * Hz[l - 1][m] - PP[l - 1][m])) / (2.0 * dz)
                              + (Hz[l][m + 1] * Hr[l][m + 1] - Hz[l][m] * Hr[l][m]) / d
                              - (n * vz[l + 1][m] - s_ * vz[l - 1][m]) / (2.0 * dz)
                              - (e * vr[l][m + 1] - c * vr[l][m]) / d)) * r[l][m]
        if which == 4:
            n, s_, e, c, a = self._axis_common(4, l, m)
            return (a + dt * (-p[l][m] * ((vz[l + 1][m] - vz[l - 1][m]) / (2.0 * dz) + (vr[l][m + 1] - vr[l][m]) / d)
                              - (n * vz[l + 1][m] - s_ * vz[l - 1][m]) / (2.0 * dz)
                              - (e * vr[l][m + 1] - c * vr[l][m]) / d)) * r[l][m]
        n, s_, e, c, a = self._axis_common(6, l, m)
        return (a + dt * ((Hr[l][m + 1] * vz[l][m + 1] - Hr[l][m] * vz[l][m]) / d
                          - (e * vr[l][m + 1] - c * vr[l][m]) / d)) * r[l][m]

    def bc_cell(self, g, l, m, ln, mn, z_l, dt):
        ph = self.ph
        if g.has_expr:
            sy = g.sym
            sy.update(gamma=self.gamma, beta=self.beta, H_z0=self.H_z0, r_0=self.r0, dz=self.dz, dr=self.dr[l],
                      r=self.r[l][m], r_z=self.rz[l][m], z=z_l, r_nb=self.r[ln][mn], r_z_nb=self.rz[ln][mn])
            for name, k in NB_SYMS:
                sy[name] = ph[k][ln][mn]
            for name, _ in COND_FIELDS:
                sy.setdefault(name, 0.0)
        for (name, k), cond, code in zip(COND_FIELDS, g.conds, g.code):
            kind = cond[0]
            if kind == "neumann":
                v = ph[k][ln][mn] + cond[1]
            elif kind == "dirichlet":
                v = cond[1]
            elif kind == "expr":
                v = float(eval(code, dict(_FUNCS), dict(g.sym)))
            else:
                v = ph[k][l][m]
            ph[k][l][m] = v
            if g.has_expr:
                g.sym[name] = v
        self.rebuild(l, m)
        kinds = [c[0] for c in g.conds]
        if kinds[0] == "axis_lf":
            self.u[0][l][m] = self.axis(0, l, m, dt)
        if kinds[1] == "axis_lf":
            self.u[1][l][m] = self.axis(1, l, m, dt)
        if kinds[7] == "axis_lf":
            self.u[4][l][m] = self.axis(4, l, m, dt)
        if kinds[4] == "axis_lf":
            self.u[6][l][m] = self.axis(6, l, m, dt)

    def face(self, f, dt):
        is_l = f in (0, 1)
        face_hi = self.Mmax if is_l else self.L - 1
        for g in self.faces[f]:
            g_lo = 0 if g.lo < 0 else g.lo
            g_hi = face_hi if g.hi < 0 else g.hi
            c_lo, c_hi = max(g_lo, 0), min(g_hi, face_hi)
            if c_lo > c_hi:
                continue
            lo, hi = c_lo + 1, c_hi + 1
            if f == 2:
                lo, hi = max(lo, 2), min(hi, self.L - 1)
                if lo > hi:
                    continue
            if is_l:
                l, ln = (1, 2) if f == 0 else (self.L, self.L - 1)
                for m in range(lo, hi + 1):
                    self.bc_cell(g, l, m, ln, m, (l - 1) * self.dz, dt)
            else:
                m, mn = (1, 2) if f == 2 else (self.Mn, self.Mn - 1)
                for l in range(lo, hi + 1):
                    self.bc_cell(g, l, m, l, mn, (l - 1) * self.dz, dt)

    def advance(self, dt):
        self.central(dt)
        self.physical(1, self.L, 2, self.Mn - 1)
        for f in (0, 3, 2, 1):  # L_LO, M_HI, M_LO, L_HI
            self.face(f, dt)
        self.physical(1, self.L, 1, self.Mn)
        for k in range(8):
            self.u0[k] = [row[:] for row in self.u[k]]

    def max_wave_speed(self):
        ph, g = self.ph, self.gamma
        mx = 0.0
        for l in range(1, self.L + 1):
            for m in range(1, self.Mn + 1):
                cs = _sqrt(_q(g * ph[P][l][m], ph[RHO][l][m]))
                h2 = ph[HZ][l][m] * ph[HZ][l][m] + ph[HR][l][m] * ph[HR][l][m] + ph[HPHI][l][m] * ph[HPHI][l][m]
                ca = _sqrt(_q(h2, ph[RHO][l][m]))
                v = _sqrt(ph[VZ][l][m] * ph[VZ][l][m] + ph[VR][l][m] * ph[VR][l][m])
                s = v + cs + ca
                mx = s if mx < s else mx
        return mx

    def compute_dt(self, dt_current):
        speed = self.max_wave_speed()
        dx = min(self.dz, self.dy)
        dt_cfl = self.cfl * dx / (speed + 1.0e-10)
        v = min(dt_cfl, dt_current * self.growth)
        return self.dt_min if v < self.dt_min else (self.dt_max if self.dt_max < v else v)
