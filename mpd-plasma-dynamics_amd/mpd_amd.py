"""ctypes binding of libmpd_lf.so, the C ABI declared in include/mpd_lf.h.

The solver mirrors the reference's Solver + Diagnostics interface
(include/solver.hpp:24-53, include/diagnostics.hpp:15-50): construct from a
config.yaml, then `advance(dt)` (Solver::Advance), `compute_dt`, `run` (the
time loop of src/main.cpp:211-247), `solution_change`, `integrals`, `outlet`.
Host arrays use the reference Fields layout: (L_max+2) x (M_max+3) float64.

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpd_lf.so")

PHYS_NAMES = ["rho", "v_z", "v_r", "v_phi", "e", "p", "P", "H_z", "H_r", "H_phi"]

MPD_OK, MPD_E_CONFIG, MPD_E_DEVICE, MPD_E_ARG, MPD_E_STATE = 0, -1, -2, -3, -4


class MpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class RunResult(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("change", C.c_double), ("steps", C.c_int),
                ("total_steps", C.c_int), ("converged", C.c_int), ("done", C.c_int), ("fast_div", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every symbol include/mpd_lf.h declares.
EXPORTS = ["mpd_create_from_file", "mpd_create_from_string", "mpd_create_with_grid", "mpd_destroy", "mpd_last_error", "mpd_dims",
           "mpd_upload_state", "mpd_reset_loop", "mpd_download", "mpd_advance", "mpd_compute_dt", "mpd_max_wave_speed",
           "mpd_solution_change", "mpd_run", "mpd_integrals", "mpd_outlet", "mpd_radial_symmetry", "mpd_grid",
           "mpd_config_check", "mpd_initial_state", "mpd_set_option", "mpd_time_steps", "mpd_probe",
           "mpd_write_vtk", "mpd_write_vtk_fields", "mpd_batch_create", "mpd_batch_destroy", "mpd_batch_size",
           "mpd_batch_instance", "mpd_batch_set_option", "mpd_batch_autotune", "mpd_batch_tile_rows",
           "mpd_batch_run", "mpd_bench_copy", "mpd_download_noalloc"]

_lib = None
_DP = C.POINTER(C.c_double)


def load():
    """Load libmpd_lf.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.mpd_create_from_file.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.mpd_create_from_string.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.mpd_create_with_grid.argtypes = [C.c_char_p, _DP, _DP, _DP, _DP, C.POINTER(vp)]
    L.mpd_destroy.argtypes = [vp]
    L.mpd_destroy.restype = None
    L.mpd_last_error.restype = C.c_char_p
    L.mpd_dims.argtypes = [vp] + [C.POINTER(C.c_int)] * 4
    L.mpd_upload_state.argtypes = [vp, C.POINTER(_DP), C.POINTER(_DP)]
    L.mpd_download.argtypes = [vp, C.POINTER(_DP), C.POINTER(_DP)]
    L.mpd_download_noalloc.argtypes = [vp, C.POINTER(_DP), C.POINTER(_DP)]
    L.mpd_reset_loop.argtypes = [vp]
    L.mpd_advance.argtypes = [vp, C.c_double]
    L.mpd_compute_dt.argtypes = [vp, C.c_double, _DP]
    L.mpd_max_wave_speed.argtypes = [vp, _DP]
    L.mpd_solution_change.argtypes = [vp, C.c_int, _DP]
    L.mpd_run.argtypes = [vp, C.c_int, C.POINTER(RunResult)]
    L.mpd_integrals.argtypes = [vp, _DP]
    L.mpd_outlet.argtypes = [vp, _DP, _DP]
    L.mpd_radial_symmetry.argtypes = [vp, _DP]
    L.mpd_grid.argtypes = [vp, _DP, _DP, _DP, _DP, _DP]
    L.mpd_config_check.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mpd_initial_state.argtypes = [C.c_char_p, C.POINTER(_DP), C.POINTER(_DP)]
    L.mpd_set_option.argtypes = [vp, C.c_char_p, C.c_double]
    L.mpd_time_steps.argtypes = [vp, C.c_int, C.c_int, _DP, _DP, _DP]
    L.mpd_batch_create.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.POINTER(vp)]
    L.mpd_batch_destroy.argtypes = [vp]
    L.mpd_batch_destroy.restype = None
    L.mpd_batch_size.argtypes = [vp, C.POINTER(C.c_int)]
    L.mpd_batch_instance.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.mpd_batch_set_option.argtypes = [vp, C.c_char_p, C.c_double]
    L.mpd_batch_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_int]
    L.mpd_batch_tile_rows.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.mpd_batch_run.argtypes = [vp, C.c_int, C.POINTER(RunResult)]
    L.mpd_bench_copy.argtypes = [C.c_longlong, C.c_int, _DP, _DP, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    for name in EXPORTS:
        if name not in ("mpd_destroy", "mpd_last_error", "mpd_batch_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != MPD_OK:
        raise MpdError(rc, load().mpd_last_error().decode(errors="replace"))


def _ptr_array(arrays):
    return (_DP * len(arrays))(*[a.ctypes.data_as(_DP) for a in arrays])


def config_check(config):
    """Host-only validation of a config text; returns (L_max, M_max)."""
    a, b = C.c_int(), C.c_int()
    _check(load().mpd_config_check(config.encode(), C.byref(a), C.byref(b)))
    return a.value, b.value


def initial_state(config):
    """Host-only step-0 state: (u0 list of 8, phys list of 10) arrays."""
    L_max, M_max = config_check(config)
    rows, cols = L_max + 2, M_max + 3
    ua = [np.zeros((rows, cols)) for _ in range(8)]
    pa = [np.zeros((rows, cols)) for _ in range(10)]
    _check(load().mpd_initial_state(config.encode(), _ptr_array(ua), _ptr_array(pa)))
    return ua, pa


def bench_copy(mib=1024, calib_launches=3):
    """Copy ceiling of the GPU's HBM (mpd_bench_copy): dict with the mean
    time of the calibration launches and the best copy rate in GB/s."""
    a, b = C.c_double(), C.c_double()
    blocks, nt = C.c_int(), C.c_int()
    _check(load().mpd_bench_copy(int(mib), int(calib_launches), C.byref(a), C.byref(b), C.byref(blocks),
                                 C.byref(nt)))
    return dict(mib=mib, calib_ms=a.value, ceiling_gbs=b.value, best_blocks=blocks.value,
                best_stores="non-temporal" if nt.value else "default")


def write_vtk_fields(config, phys, path, binary=True):
    """Host-only: the solver's VTK frame of 10 host physical arrays."""
    pa = [np.ascontiguousarray(a, dtype=np.float64) for a in phys]
    _check(load().mpd_write_vtk_fields(config.encode(), _ptr_array(pa), str(path).encode(), 1 if binary else 0))


class Solver:
    """Device-resident LF solver built from a config (path or YAML text)."""

    def __init__(self, config=None, *, path=None, grid_rows=None, _handle=None, _owner=None):
        """grid_rows: optional (r_in, r_out, dr_in_dz, dr_out_dz), each L_max + 2
        values at z = (l - 1) / L_max, replacing the config's geometry."""
        L = load()
        h = C.c_void_p()
        if _handle is not None:
            h = _handle
        elif grid_rows is not None:
            rows = [np.ascontiguousarray(a, dtype=np.float64) for a in grid_rows]
            _check(L.mpd_create_with_grid(config.encode(), *[a.ctypes.data_as(_DP) for a in rows], C.byref(h)))
        elif path is not None:
            _check(L.mpd_create_from_file(str(path).encode(), C.byref(h)))
        else:
            _check(L.mpd_create_from_string(config.encode(), C.byref(h)))
        self._h = h
        self._owner = _owner  # a Batch keeps its instances alive
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(L.mpd_dims(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        self.L_max, self.M_max, self.rows, self.cols = a.value, b.value, c.value, d.value

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                load().mpd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _new(self, n):
        return [np.zeros((self.rows, self.cols), dtype=np.float64) for _ in range(n)]

    def download(self, u0=True, phys=True):
        """Returns (u0 list of 8 arrays | None, phys list of 10 arrays | None)."""
        ua = self._new(8) if u0 else None
        pa = self._new(10) if phys else None
        _check(load().mpd_download(self._h, _ptr_array(ua) if ua else None, _ptr_array(pa) if pa else None))
        return ua, pa

    def upload(self, u0, phys=None):
        ua = [np.ascontiguousarray(a, dtype=np.float64) for a in u0]
        pa = [np.ascontiguousarray(a, dtype=np.float64) for a in phys] if phys is not None else None
        for a in ua + (pa or []):
            if a.shape != (self.rows, self.cols):
                raise ValueError(f"array shape {a.shape} != {(self.rows, self.cols)}")
        _check(load().mpd_upload_state(self._h, _ptr_array(ua), _ptr_array(pa) if pa else None))

    def reset_loop(self):
        """Loop variables of a fresh run: t = 0, dt = time.dt, no steps."""
        _check(load().mpd_reset_loop(self._h))

    def cold_restart(self, config):
        """Back to the config's initial condition and a fresh time loop (the
        same run as a newly created solver's)."""
        u0, ph = initial_state(config)
        self.upload(u0, ph)
        self.reset_loop()
        self.solution_change(save_prev=True)

    def advance(self, dt):
        _check(load().mpd_advance(self._h, float(dt)))

    def compute_dt(self, dt_current):
        v = C.c_double()
        _check(load().mpd_compute_dt(self._h, float(dt_current), C.byref(v)))
        return v.value

    def max_wave_speed(self):
        v = C.c_double()
        _check(load().mpd_max_wave_speed(self._h, C.byref(v)))
        return v.value

    def solution_change(self, save_prev=True):
        v = C.c_double()
        _check(load().mpd_solution_change(self._h, 1 if save_prev else 0, C.byref(v)))
        return v.value

    def run(self, max_steps=0):
        r = RunResult()
        _check(load().mpd_run(self._h, int(max_steps), C.byref(r)))
        return r.as_dict()

    def integrals(self):
        out = (C.c_double * 7)()
        _check(load().mpd_integrals(self._h, out))
        return np.array(out[:])

    def outlet(self):
        a, b = C.c_double(), C.c_double()
        _check(load().mpd_outlet(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def radial_symmetry(self):
        v = C.c_double()
        _check(load().mpd_radial_symmetry(self._h, C.byref(v)))
        return v.value

    def grid(self):
        r = np.zeros((self.rows, self.cols))
        rz = np.zeros((self.rows, self.cols))
        R = np.zeros(self.rows)
        dr = np.zeros(self.rows)
        r0 = C.c_double()
        _check(load().mpd_grid(self._h, r.ctypes.data_as(_DP), rz.ctypes.data_as(_DP), R.ctypes.data_as(_DP),
                               dr.ctypes.data_as(_DP), C.byref(r0)))
        return dict(r=r, r_z=rz, R=R, dr=dr, r_0=r0.value)

    def probe(self, l, m):
        """Physical values (10, Fields member order) of cell (l, m)."""
        out = (C.c_double * 10)()
        _check(load().mpd_probe(self._h, int(l), int(m), out))
        return np.array(out[:])

    def write_vtk(self, path):
        _check(load().mpd_write_vtk(self._h, str(path).encode()))

    def set_option(self, key, value):
        _check(load().mpd_set_option(self._h, key.encode(), float(value)))

    def time_steps(self, steps, warmup=0):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(load().mpd_time_steps(self._h, int(steps), int(warmup), C.byref(a), C.byref(b), C.byref(c)))
        return dict(total_ms=a.value, stencil_ms=b.value, bc_ms=c.value)


class Batch:
    """Many independent solver instances (configs of any grid size) advanced
    together by the batched kernels; `instance(i)` is a Solver view of one."""

    def __init__(self, configs):
        L = load()
        texts = [c.encode() for c in configs]
        arr = (C.c_char_p * len(texts))(*texts)
        h = C.c_void_p()
        _check(L.mpd_batch_create(arr, len(texts), C.byref(h)))
        self._h = h
        self.n = len(texts)
        self._views = {}

    def close(self):
        if getattr(self, "_h", None):
            for v in self._views.values():
                v._h = None
            self._views = {}
            load().mpd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def instance(self, i):
        if i not in self._views or self._views[i]._h is None:
            h = C.c_void_p()
            _check(load().mpd_batch_instance(self._h, int(i), C.byref(h)))
            self._views[i] = Solver(_handle=h, _owner=self)
        return self._views[i]

    def set_option(self, key, value):
        _check(load().mpd_batch_set_option(self._h, key.encode(), float(value)))

    def autotune(self, reps=20):
        buf = C.create_string_buffer(4096)
        _check(load().mpd_batch_autotune(self._h, int(reps), buf, len(buf)))
        return buf.value.decode()

    def tile_rows(self, i):
        v = C.c_int()
        _check(load().mpd_batch_tile_rows(self._h, int(i), C.byref(v)))
        return v.value

    def warm_start(self, i, u0):
        """Replace instance i's state by u0 (8 arrays, physical = f(u0)) and
        take it as the convergence snapshot."""
        s = self.instance(i)
        s.upload(u0)
        s.solution_change(save_prev=True)

    def run(self, max_steps=0):
        res = (RunResult * self.n)()
        _check(load().mpd_batch_run(self._h, int(max_steps), res))
        return [r.as_dict() for r in res]
