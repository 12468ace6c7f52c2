"""ctypes binding of the CPU oracle (oracle/liblf_oracle.so) and a driver for the
reference harness (oracle/_ref/ref_harness).  Test infrastructure only: the
product never imports this module."""
import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liblf_oracle.so")
REF_HARNESS = os.path.join(ORACLE_DIR, "_ref", "ref_harness")

NEUMANN, DIRICHLET, EXPRESSION, AXIS_LF = 0, 1, 2, 3
GEOM_STRAIGHT, GEOM_CONDIV, GEOM_CATHODE_COSINE, GEOM_ROWS = 0, 1, 2, 3

# physical ids (include/fields.hpp member order)
RHO, VZ, VR, VPHI, E, P, PT, HZ, HR, HPHI = range(10)
PHYS_NAMES = ["rho", "v_z", "v_r", "v_phi", "e", "p", "P", "H_z", "H_r", "H_phi"]
# BC / IC field order (evaluation order)
BC_FIELDS = ["rho", "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"]


class Cond(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double), ("expr", C.c_char_p)]


class Seg(C.Structure):
    _fields_ = [("lo", C.c_int), ("hi", C.c_int), ("c", Cond * 8)]


class Params(C.Structure):
    _fields_ = [
        ("L_max", C.c_int), ("M_max", C.c_int),
        ("gamma", C.c_double), ("beta", C.c_double), ("H_z0", C.c_double),
        ("T", C.c_double), ("dt", C.c_double),
        ("adaptive", C.c_int),
        ("cfl", C.c_double), ("growth", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("conv_threshold", C.c_double), ("check_frequency", C.c_int),
        ("geom_type", C.c_int), ("geom_p", C.c_double * 8),
        ("ic_expr", C.c_char_p * 8),
        ("n_ic_vars", C.c_int), ("ic_var_names", C.POINTER(C.c_char_p)),
        ("ic_var_values", C.POINTER(C.c_double)),
        ("ic_noise_amp", C.c_double), ("ic_noise_seed", C.c_ulonglong),
        ("nseg", C.c_int * 4), ("seg", C.POINTER(Seg) * 4),
        ("geom_rows", C.POINTER(C.c_double) * 4),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liblf_oracle.so"])
        L = C.CDLL(ORACLE_SO)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), C.c_char_p, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_rows.argtypes = [C.c_void_p]
        L.orc_cols.argtypes = [C.c_void_p]
        L.orc_array.restype = C.POINTER(C.c_double)
        L.orc_array.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_r0.restype = C.c_double
        L.orc_r0.argtypes = [C.c_void_p]
        L.orc_advance.argtypes = [C.c_void_p, C.c_double]
        for fn in ("orc_max_wave_speed", "orc_solution_change", "orc_mass_flux", "orc_thrust",
                   "orc_radial_symmetry"):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_compute_dt.restype = C.c_double
        L.orc_compute_dt.argtypes = [C.c_void_p, C.c_double]
        L.orc_save_prev.argtypes = [C.c_void_p]
        L.orc_update_physical.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_init_conservative.argtypes = [C.c_void_p]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_integrals.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_axis_lf.restype = C.c_double
        L.orc_axis_lf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_noise_u01.restype = C.c_double
        L.orc_noise_u01.argtypes = [C.c_ulonglong, C.c_ulonglong]
        _lib = L
    return _lib


def _cond(spec):
    """spec: ('neumann', off) | ('dirichlet', v) | ('expr', 'text') | ('axis_lf',)"""
    c = Cond()
    kind = spec[0]
    if kind == "neumann":
        c.type, c.value = NEUMANN, float(spec[1]) if len(spec) > 1 else 0.0
    elif kind == "dirichlet":
        c.type, c.value = DIRICHLET, float(spec[1]) if len(spec) > 1 else 0.0
    elif kind == "expr":
        c.type, c.expr = EXPRESSION, spec[1].encode()
    elif kind == "axis_lf":
        c.type = AXIS_LF
    else:
        raise ValueError(kind)
    return c


class Oracle:
    """One oracle solver.  case: dict with keys L, M, gamma, beta, H_z0, T, dt,
    adaptive, cfl, growth, dt_min, dt_max, conv_threshold, check_frequency,
    geom (type, params), ic (dict field->expr), ic_vars, noise_amp, noise_seed,
    faces {l_lo, l_hi, m_lo, m_hi: [ {range: (lo, hi), field: spec} ]}."""

    def __init__(self, case):
        L = lib()
        p = Params()
        p.L_max, p.M_max = case["L"], case["M"]
        p.gamma = case.get("gamma", 1.67)
        p.beta = case.get("beta", 0.05)
        p.H_z0 = case.get("H_z0", 0.25)
        p.T = case.get("T", 1e30)
        p.dt = case.get("dt", 2.5e-5)
        p.adaptive = 1 if case.get("adaptive", False) else 0
        p.cfl = case.get("cfl", 0.5)
        p.growth = case.get("growth", 1.1)
        p.dt_min = case.get("dt_min", 1e-9)
        p.dt_max = case.get("dt_max", 1e-3)
        p.conv_threshold = case.get("conv_threshold", 0.0)
        p.check_frequency = case.get("check_frequency", 100)
        gt, gp = case.get("geom", (GEOM_STRAIGHT, [0.2, 0.8]))
        self._keep = []
        if case.get("geom_rows") is not None:
            gt = GEOM_ROWS
            rows = [np.ascontiguousarray(a, dtype=np.float64) for a in case["geom_rows"]]
            assert all(a.shape == (case["L"] + 2,) for a in rows)
            self._keep += rows
            for i, a in enumerate(rows):
                p.geom_rows[i] = a.ctypes.data_as(C.POINTER(C.c_double))
        else:
            for i, v in enumerate(gp):
                p.geom_p[i] = v
        p.geom_type = gt
        ic = case.get("ic", {})
        for i, f in enumerate(BC_FIELDS):
            if f in ic:
                b = str(ic[f]).encode()
                self._keep.append(b)
                p.ic_expr[i] = b
        vars_ = case.get("ic_vars", {})
        if vars_:
            names = (C.c_char_p * len(vars_))(*[k.encode() for k in vars_])
            vals = (C.c_double * len(vars_))(*[float(v) for v in vars_.values()])
            self._keep += [names, vals]
            p.n_ic_vars = len(vars_)
            p.ic_var_names = C.cast(names, C.POINTER(C.c_char_p))
            p.ic_var_values = C.cast(vals, C.POINTER(C.c_double))
        p.ic_noise_amp = case.get("noise_amp", 0.0)
        p.ic_noise_seed = case.get("noise_seed", 0)
        faces = case.get("faces", {})
        for fi, fname in enumerate(["l_lo", "l_hi", "m_lo", "m_hi"]):
            segs = faces.get(fname, [])
            arr = (Seg * max(1, len(segs)))()
            for si, sg in enumerate(segs):
                lo, hi = sg.get("range", (-1, -1))
                arr[si].lo, arr[si].hi = lo, hi
                for k, f in enumerate(BC_FIELDS):
                    arr[si].c[k] = _cond(sg.get(f, ("neumann", 0.0)))
            self._keep.append(arr)
            p.nseg[fi] = len(segs)
            p.seg[fi] = C.cast(arr, C.POINTER(Seg))
        self._p = p
        err = C.create_string_buffer(512)
        h = L.orc_create(C.byref(p), err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        self.h = h
        self.rows = L.orc_rows(h)
        self.cols = L.orc_cols(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def array(self, kind, k=0):
        ptr = lib().orc_array(self.h, kind, k)
        n = self.rows if kind in (6, 7) else self.rows * self.cols
        a = np.ctypeslib.as_array(ptr, shape=(n,))
        return a if kind in (6, 7) else a.reshape(self.rows, self.cols)

    def u0(self, k):
        return self.array(0, k)

    def u(self, k):
        return self.array(1, k)

    def phys(self, k):
        return self.array(2, k)

    def prev(self, k):
        return self.array(3, k)

    @property
    def r(self):
        return self.array(4)

    @property
    def r_z(self):
        return self.array(5)

    @property
    def dr(self):
        return self.array(6)

    @property
    def R(self):
        return self.array(7)

    @property
    def r0(self):
        return lib().orc_r0(self.h)

    def advance(self, dt):
        lib().orc_advance(self.h, dt)

    def max_wave_speed(self):
        return lib().orc_max_wave_speed(self.h)

    def compute_dt(self, dt):
        return lib().orc_compute_dt(self.h, dt)

    def solution_change(self):
        return lib().orc_solution_change(self.h)

    def save_prev(self):
        lib().orc_save_prev(self.h)

    def update_physical(self, l_lo, l_hi, m_lo, m_hi):
        lib().orc_update_physical(self.h, l_lo, l_hi, m_lo, m_hi)

    def init_conservative(self):
        lib().orc_init_conservative(self.h)

    def run(self, max_steps, t=0.0, dt=None):
        tt = C.c_double(t)
        dd = C.c_double(self._p.dt if dt is None else dt)
        conv = C.c_int(0)
        ch = C.c_double(0)
        n = lib().orc_run(self.h, max_steps, C.byref(tt), C.byref(dd), C.byref(conv), C.byref(ch))
        return dict(steps=n, t=tt.value, dt=dd.value, converged=bool(conv.value), change=ch.value)

    def axis_lf(self, k, l, m, dt):
        """AxisLfU(k+1) at (l, m) on the current arrays (k = 0, 1, 4, 6)."""
        return lib().orc_axis_lf(self.h, k, l, m, dt)

    def integrals(self):
        out = (C.c_double * 7)()
        lib().orc_integrals(self.h, out)
        return np.array(out[:])

    def mass_flux(self):
        return lib().orc_mass_flux(self.h)

    def thrust(self):
        return lib().orc_thrust(self.h)

    def radial_symmetry(self):
        return lib().orc_radial_symmetry(self.h)


def ref_harness_available():
    return os.path.exists(REF_HARNESS)


def run_ref_harness(L, M, geom, gp, gamma, steps, dt, adaptive, cfl, growth, dt_min, dt_max,
                    phys, u0):
    """Run the reference-compiled harness on the state (phys: 10 arrays, u0: 8
    arrays of shape (L+2, M+3)).  Returns a dict of its outputs."""
    rows, cols = L + 2, M + 3
    gp = list(gp) + [0.0] * (5 - len(gp))
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for a in list(phys) + list(u0):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        args = [REF_HARNESS, str(L), str(M), str(geom)] + [repr(float(v)) for v in gp[:5]] + [
            repr(float(gamma)), str(steps), repr(float(dt)), str(int(adaptive)), repr(float(cfl)),
            repr(float(growth)), repr(float(dt_min)), repr(float(dt_max)), fin, fout]
        env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "4"))
        pr = subprocess.run(args, check=True, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
        raw = np.fromfile(fout, dtype=np.float64)
    loop_s = [float(x.split()[1]) for x in pr.stderr.splitlines() if x.startswith("loop_seconds")]
    n2 = rows * cols
    o = {}
    pos = 0

    def take(n):
        nonlocal pos
        v = raw[pos:pos + n]
        pos += n
        return v

    o["r"] = take(n2).reshape(rows, cols)
    o["r_z"] = take(n2).reshape(rows, cols)
    o["R"] = take(rows)
    o["dr"] = take(rows)
    o["r0"] = take(1)[0]
    o["speed0"] = take(1)[0]
    o["integrals0"] = take(7)
    o["u0"] = [take(n2).reshape(rows, cols) for _ in range(8)]
    o["phys"] = [take(n2).reshape(rows, cols) for _ in range(10)]
    o["dts"] = take(steps)
    o["dt_next"] = take(1)[0]
    o["change"] = take(1)[0]
    o["speed"] = take(1)[0]
    o["integrals"] = take(7)
    o["mass_flux"] = take(1)[0]
    o["thrust"] = take(1)[0]
    o["loop_seconds"] = loop_s[0] if loop_s else None
    assert pos == raw.size, (pos, raw.size)
    return o
