"""GPU test of mpd_download_noalloc, the download that allocates no device
memory (its staging goes through the library's per-device `scratch` array),
through the C ABI: every array equal bit for bit to mpd_download's, on calls
in a row with a smaller grid second, on a grid whose rows take several
scratch chunks, before the first step (physical arrays of the IC), and the
clear error for a row wider than the scratch."""
import ctypes as C

import numpy as np
import pytest

import cases
from parity import same_bits

pytestmark = pytest.mark.gpu


def _noalloc(mpd):
    L = mpd.load()
    f = L.mpd_download_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(mpd._DP), C.POINTER(mpd._DP)]
    f.restype = C.c_int
    return f


def _download_both(mpd, sol, with_u0=True):
    f = _noalloc(mpd)
    ua = [np.full((sol.rows, sol.cols), np.nan) for _ in range(8)] if with_u0 else None
    pa = [np.full((sol.rows, sol.cols), np.nan) for _ in range(10)]
    rc = f(sol._h, mpd._ptr_array(ua) if ua else None, mpd._ptr_array(pa))
    assert rc == 0, mpd.load().mpd_last_error()
    ub, pb = sol.download(u0=with_u0, phys=True)
    return ua, pa, ub, pb


def _check(mpd, sol, label, with_u0=True):
    ua, pa, ub, pb = _download_both(mpd, sol, with_u0)
    if with_u0:
        for k in range(8):
            assert same_bits(ua[k], ub[k]), f"{label}: u0_{k + 1}"
    for k in range(10):
        assert same_bits(pa[k], pb[k]), f"{label}: {mpd.PHYS_NAMES[k]}"


def test_noalloc_download_matches_download(mpd):
    big = cases.channel_config(200, 80, geometry="condiv", noise=1e-3, max_steps=60)
    small = cases.channel_config(40, 16, max_steps=25)
    with mpd.Solver(big) as a:
        _check(mpd, a, "200x80 before the first step")  # physical arrays of the IC
        a.run()
        _check(mpd, a, "200x80 after 60 steps")
        with mpd.Solver(small) as b:
            b.run()
            _check(mpd, b, "40x16 after 25 steps (second, smaller call)")
            _check(mpd, b, "40x16 physical arrays only", with_u0=False)
        _check(mpd, a, "200x80 again")


def test_noalloc_download_in_row_chunks(mpd):
    """At M_max = 2560 one row of the 10 physical fields is 26,880 doubles, so
    the 16 MiB scratch holds 78 rows: 302 rows take four chunks, the last
    partial."""
    text = cases.channel_config(300, 2560, geometry="condiv", noise=1e-3, max_steps=3)
    with mpd.Solver(text) as sol:
        sol.run()
        _check(mpd, sol, "300x2560")


def test_noalloc_download_rejects_rows_wider_than_scratch(mpd):
    text = cases.channel_config(3, 210000, max_steps=1)
    with mpd.Solver(text) as sol:
        f = _noalloc(mpd)
        pa = [np.zeros((sol.rows, sol.cols)) for _ in range(10)]
        rc = f(sol._h, None, mpd._ptr_array(pa))
        assert rc == mpd.MPD_E_ARG
        assert b"no-allocation scratch" in mpd.load().mpd_last_error()
        # The conserved arrays alone need no staging.
        ua = [np.zeros((sol.rows, sol.cols)) for _ in range(8)]
        assert f(sol._h, mpd._ptr_array(ua), None) == 0
        ub, _ = sol.download(phys=False)
        assert all(same_bits(x, y) for x, y in zip(ua, ub))
