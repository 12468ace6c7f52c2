"""The host-side setup run from several threads at once under
ThreadSanitizer (CPU only): a batch builds its instances' configs, geometry
registry entries and geometries on a thread pool, so that code must be free
of data races.  The check program (host_threads_check.cpp) is compiled with
g++ -fsanitize=thread together with the library's host sources."""
import glob
import os
import subprocess

import pytest

import cases

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "mpd-plasma-dynamics_amd", "host")


def test_host_setup_has_no_data_races(tmp_path):
    exe = tmp_path / "host_threads_check"
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-ffp-contract=off",
           "-fno-builtin-sin", "-fno-builtin-cos", "-o", str(exe),
           os.path.join(REPO, "tests", "host_threads_check.cpp")] + srcs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "sanitizer" in (r.stderr or "").lower() and "not" in r.stderr.lower():
        pytest.skip("ThreadSanitizer runtime not available: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-3000:]
    cfg = tmp_path / "config.yaml"
    cfg.write_text(cases.channel_config(24, 10, geometry="condiv", max_steps=5))
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([str(exe), str(cfg)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    assert "ok" in r.stdout
