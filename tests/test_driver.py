"""The simulation driver mpd_lf_run end to end on the GPU: the shipped
thruster case's physics and boundary program (cosine cathode geometry,
3-segment M_LO with axis_lf, expression BCs) at a reduced grid, run through
the reference's time loop with convergence reports, VTK frames and
checkpoint lines (src/main.cpp:22-317).  Every number the driver prints and
the last frame's fields are checked against the oracle run in the same
steps."""
import os
import re
import subprocess

import numpy as np
import pytest

import _oracle as O
import cases
from test_vtk import _parse_legacy

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "mpd-plasma-dynamics_amd", "mpd_lf_run")

STEPS, CHECK, VTK = 2000, 500, 1000


def _config(L, M):
    text = cases.thruster_config(L, M, max_steps=STEPS)
    return text + f"""
convergence:
  threshold: 1.0e-30
  check_frequency: {CHECK}
output:
  directory: out
  run_name: drv
  vtk_step: {VTK}
"""


def _close(printed, value, rel):
    """A number printed with a %e format agrees with value to the print's precision."""
    p = float(printed)
    return p == value or abs(p - value) <= rel * abs(value)


@pytest.mark.gpu
def test_driver_thruster_end_to_end(tmp_path):
    assert os.path.exists(DRIVER), "build mpd_lf_run first"
    L, M = 80, 40
    text = _config(L, M)
    (tmp_path / "cfg.yaml").write_text(text)
    print(f"driver: thruster {L}x{M}, {STEPS} steps", flush=True)
    p = subprocess.run([DRIVER, "cfg.yaml"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    out = p.stdout.splitlines()

    # Header and verification lines (src/main.cpp:42-70, 76-79, 117-136).
    assert out[0] == "Config file         : cfg.yaml"
    assert out[1] == f"Grid                : {L} x {M}  (L x M)"
    assert out[2] == "Initial dt / T end  : 1.000000e-07 / 6.0000"
    assert out[3] == "MPI ranks           : 1  (1 x 1 Cartesian)"
    assert out[4] == "Geometry            : short_cathode_cosine"
    assert out[5] == "Adaptive dt         : ON  (CFL=0.45, growth=1.10, dt_min=1.0e-10, dt_max=5.0e-05)"
    assert out[6] == f"Convergence check   : threshold = 1.00e-30, every {CHECK} steps"
    assert out[7] == f"VTK output every    : {VTK} steps"
    assert "[Verification] Ghost exchange  : PASS" in out
    assert "IC                  : expression" in out

    case = cases.oracle_case(text)
    o = O.Oracle(case)
    o.save_prev()  # Fields::SavePrev at start-up with convergence on (src/main.cpp:150-152)
    sym = [l for l in out if l.startswith("[Verification] Radial symmetry")][0]
    assert _close(re.search(r"= ([0-9.e+-]+)", sym).group(1), o.radial_symmetry(), 1e-3)
    want = "field is radially uniform" if o.radial_symmetry() < 1.0e-10 else "radial variation present"
    assert sym.endswith(f"({want})")

    rows = [l.split() for l in out if re.match(r"^  \d+ ", l)]
    ref = o.integrals()
    assert rows[0][0] == "0" and all(_close(v, w, 1e-6) for v, w in zip(rows[0][2:], ref))

    # Loop: oracle in chunks of CHECK steps, compared with each report.
    conv = [l for l in out if l.startswith("Step ")]
    probe = [l.split() for l in out if re.match(r"^\d+\.\d{6} +\d\.\d{6}e[+-]\d\d ", l)]
    t, dt = 0.0, case["dt"]
    n_integral = 1
    for k in range(1, STEPS // CHECK + 1):
        res = o.run(CHECK, t=t, dt=dt)
        assert res["steps"] == CHECK
        t, dt = res["t"], res["dt"]
        step = k * CHECK
        m = re.match(r"Step (\d+), t=(\S+), dt=(\S+), relative change: (\S+)$", conv[k - 1])
        assert int(m.group(1)) == step
        assert m.group(2) == f"{t:.6f}" and m.group(3) == f"{dt:.6e}" and _close(m.group(4), res["change"], 1e-6)
        if step % 1000 == 0:
            pr = probe[step // 1000 - 1]
            ph = [o.phys(kk)[21, 41] for kk in (0, 1, 3, 4, 9)]
            assert pr[0] == f"{t:.6f}" and pr[1] == f"{dt:.6e}"
            assert pr[2:] == [f"{v:.6f}" for v in ph]
            q = o.integrals()
            row = rows[n_integral]
            n_integral += 1
            assert int(row[0]) == step and all(_close(v, w, 1e-6) for v, w in zip(row[2:], q))

    assert any(l.startswith(f"Calculation time : ") and l.endswith(f"({STEPS} steps)") for l in out)
    drift = out.index(f"[Verification] Drift from step-0 reference at step {STEPS} (t={t:.5f})")
    names = ["mass", "mom_z", "mom_r", "mom_phi", "energy", "mag_z", "mag_phi"]
    fin = o.integrals()
    for i, nm in enumerate(names):
        f = out[drift + 3 + i].split()
        assert f[0] == nm and _close(f[1], ref[i], 1e-8) and _close(f[2], fin[i], 1e-8)
    mf = [l for l in out if l.startswith("Mass flux : ")][0]
    th = [l for l in out if l.startswith("Thrust    : ")][0]
    assert _close(mf.split(":")[1], o.mass_flux(), 1e-6) and _close(th.split(":")[1], o.thrust(), 1e-6)

    # Frames: step 0, every VTK steps, and the final one; the last holds the
    # oracle's physical fields bit for bit.
    run_dirs = os.listdir(tmp_path / "out")
    assert len(run_dirs) == 1 and re.match(r"drv_\d\d-\d\d-\d{4}_\d\d:\d\d:\d\d:\d{3}$", run_dirs[0])
    d = tmp_path / "out" / run_dirs[0]
    assert sorted(os.listdir(d)) == ["step_0000.vtk", "step_1000.vtk", "step_2000.vtk"]
    _, ni, nj, _, arr = _parse_legacy(d / "step_2000.vtk")
    li = np.minimum(np.arange(ni), L - 1)

    def g(k):
        return o.phys(k)[li + 1][:, 1:M + 2].T

    assert np.array_equal(arr["Rho"][:, :, 0], g(0))
    assert np.array_equal(arr["Energy"][:, :, 0], g(4))
    assert np.array_equal(arr["Velocity"], np.stack([g(1), g(2), g(3)], axis=-1))
    assert np.array_equal(arr["MagneticField"], np.stack([g(7), g(8), g(9)], axis=-1))


SWEEP = os.path.join(REPO, "mpd-plasma-dynamics_amd", "mpd_sweep_run")


@pytest.mark.gpu
def test_sweep_driver_cases_match_oracle(tmp_path):
    """mpd_sweep_run on three configs of different size and boundary program
    (one stops on convergence, two on their step caps): every case's printed
    loop end and outlet diagnostics, and its final VTK frame, are the
    oracle's run of that config alone."""
    assert os.path.exists(SWEEP), "build mpd_sweep_run first"
    texts = [cases.channel_config(40, 16, max_steps=60),
             cases.thruster_config(48, 24, max_steps=40),
             cases.channel_config(30, 12, max_steps=5000, rho="1.3", threshold=2e-3, check_frequency=10)]
    names = []
    for i, t in enumerate(texts):
        (tmp_path / f"c{i}.yaml").write_text(t)
        names.append(f"c{i}.yaml")
    print("sweep driver: 3 cases", flush=True)
    p = subprocess.run([SWEEP, "--out", "frames", *names], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    out = p.stdout.splitlines()
    assert out[0] == "[sweep] 3 cases"
    rows = {int(l.split()[0]): l.split() for l in out if re.match(r"^\d+ +c\d\.yaml ", l)}
    assert sorted(rows) == [0, 1, 2]
    for i, t in enumerate(texts):
        case = cases.oracle_case(t)
        o = O.Oracle(case)
        res = o.run(case["max_steps"])
        r = rows[i]
        assert r[1] == names[i] and r[2] == f"{case['L']}x{case['M']}"
        assert int(r[3]) == res["steps"], (i, r, res)
        assert _close(r[4], res["t"], 1e-6) and _close(r[5], res["dt"], 1e-6)
        assert r[7] == ("yes" if res["converged"] else "no")
        assert _close(r[8], o.mass_flux(), 1e-6) and _close(r[9], o.thrust(), 1e-6)
        _, ni, nj, _, arr = _parse_legacy(tmp_path / "frames" / f"case_{i:04d}.vtk")
        li = np.minimum(np.arange(ni), case["L"] - 1)
        assert np.array_equal(arr["Rho"][:, :, 0], o.phys(0)[li + 1][:, 1:case["M"] + 2].T)
    assert rows[2][7] == "yes" and rows[0][7] == "no"
