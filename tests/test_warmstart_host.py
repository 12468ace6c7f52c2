"""Host-side (CPU) tests of the warm-start pipeline's pure logic: the
fallback rule, the admissibility screen of predictions, the target
normalization, the features and the conserved-state rebuild."""
import numpy as np

import surrogate as SG
import sweep
import warmstart as WS


def test_stall_rule():
    # Ended without converging (e.g. t became NaN): fall back.
    assert WS.stalled(True, False, 1e-3, 1e-3, 1e-6)
    assert not WS.stalled(True, True, 5e-9, 1e-7, 1e-6)
    # Running: non-finite change, or above the threshold and not falling.
    assert WS.stalled(False, False, float("nan"), 1e-4, 1e-6)
    assert WS.stalled(False, False, 2e-4, 1e-4, 1e-6)
    assert WS.stalled(False, False, 1e-4, 1e-4, 1e-6)
    assert not WS.stalled(False, False, 0.9e-4, 1e-4, 1e-6)
    assert not WS.stalled(False, False, 5e-7, 1e-7, 1e-6)


def test_step_budget():
    """The warm-start step budget: 1.5 x the slowest training case's cold
    step count (none without training cases)."""
    assert WS.step_budget([1000, 4000, 2500]) == 6000
    assert WS.step_budget(np.array([7]), factor=2.0) == 14
    assert WS.step_budget([]) is None


def test_admissible():
    f = np.ones((4, 3, 8))
    assert WS.admissible(f)
    g = f.copy(); g[1, 1, 0] = -1e-3
    assert not WS.admissible(g)
    g = f.copy(); g[0, 2, 4] = 0.0
    assert not WS.admissible(g)
    g = f.copy(); g[3, 0, 6] = np.inf
    assert not WS.admissible(g)
    g = f.copy(); g[..., 2] = -5.0  # velocities may be negative
    assert WS.admissible(g)


def test_normalizer_round_trip_and_dead_channels():
    rng = np.random.default_rng(0)
    r = rng.normal(size=(10, 6, 5, 8))
    r[..., 3] = 0.0  # a field the sweep never excites
    n = WS.Normalizer(r)
    assert not n.live[3] and n.live[[0, 1, 2, 4, 5, 6, 7]].all()
    e = n.encode(r)
    assert np.all(e[..., 3] == 0.0)
    np.testing.assert_allclose(n.decode(e), r, rtol=1e-12, atol=1e-12)
    # Whatever the network outputs in a dead channel, the prediction is exact.
    out = e.copy(); out[..., 3] = 7.0
    assert np.all(n.decode(out)[..., 3] == 0.0)


def test_grid_radii_closed_form():
    p = np.array([[1.0, 0.1, 0.2, 0.05, 0.1]])
    r = WS.grid_radii(p, 10, 4)
    assert r.shape == (1, 12, 7)
    # Interior cell (l=1, m=1): z = 0, frac = 0 -> r_in(0) = 0.2.
    assert r[0, 1, 1] == 0.2
    # (l = L, m = M + 1): z = 0.9, frac = 1 -> r_out(0.9) = 0.8 - 0.1 * 0.81.
    assert abs(r[0, 10, 5] - (0.8 - 0.1 * 0.81)) < 1e-15


def test_u0_from_fields_is_rebuild_from_physical():
    """RebuildUFromPhysical (include/physics_utils.hpp:15-24): u1 = rho r,
    u2..u5 = rho (v_z, v_r, v_phi, e) r, u6 = H_phi, u7, u8 = (H_z, H_r) r."""
    rng = np.random.default_rng(1)
    L, M = 6, 3
    f = rng.uniform(0.5, 2.0, size=(L, M + 1, 8))
    r = rng.uniform(0.2, 0.8, size=(L + 2, M + 3))
    u = SG.u0_from_fields(f, r, L + 2, M + 3)
    rr = r[1:L + 1, 1:M + 2]
    rho, vz, vr, vphi, e, hz, hr, hphi = (f[..., k] for k in range(8))
    for k, v in enumerate([rho * rr, rho * vz * rr, rho * vr * rr, rho * vphi * rr, rho * e * rr, hphi,
                           hz * rr, hr * rr]):
        assert np.array_equal(u[k][1:L + 1, 1:M + 2], v)
        assert np.all(u[k][0] == 0) and np.all(u[k][:, 0] == 0) and np.all(u[k][L + 1] == 0)


def test_features_and_baseline():
    p = sweep.sample_params(3, 4)
    r = WS.grid_radii(p, 8, 3)
    x = SG.features(p, r, 8, 3)
    assert x.shape == (3, 8, 4, 8) and x.dtype == np.float32
    assert np.all(np.abs(x[..., :5]) <= 1.0)
    b = SG.baseline_fields(p, r, 8, 3, WS.GAMMA)
    np.testing.assert_allclose(b[..., 7], p[:, 2, None, None] / r[:, 1:9, 1:5])
    np.testing.assert_allclose(b[..., 4] * (WS.GAMMA - 1) * b[..., 0], 0.5)


def test_case_config_parses():
    import mpd_amd
    p = sweep.sample_params(2, 9)
    for q in p:
        assert mpd_amd.config_check(sweep.case_config(q, 20, 8)) == (20, 8)


def test_padding_for_the_network_levels():
    """Grids whose sides are not multiples of 8 run on the padded size:
    edge replication up, cropping back is exact."""
    assert SG.padded(25) == 32 and SG.padded(24) == 24 and SG.padded(161) == 168 and SG.padded(321) == 328
    a = np.arange(2 * 5 * 3 * 2, dtype=np.float64).reshape(2, 5, 3, 2)
    p = SG.pad_grid(a, 8, 8)
    assert p.shape == (2, 8, 8, 2)
    assert np.array_equal(p[:, :5, :3], a)
    assert np.array_equal(p[:, 5:, :3], np.repeat(a[:, 4:5], 3, axis=1))
    assert np.array_equal(p[:, :5, 3:], np.repeat(a[:, :, 2:3], 5, axis=2))
    assert np.all(p[:, 5:, 3:] == a[:, 4:5, 2:3])
    assert SG.pad_grid(a, 5, 3) is a


def test_resample_fields_bilinear_exact():
    """A field bilinear in (z, frac) is reproduced by the interpolation onto a
    finer grid wherever the target lies inside the source's z range (constant
    beyond its last row); the same grid is the identity."""
    L0, M0, L1, M1 = 10, 4, 20, 8
    z0 = np.arange(L0) / L0
    y0 = np.arange(M0 + 1) / M0
    f = np.zeros((2, L0, M0 + 1, 8))
    for k in range(8):
        f[0, ..., k] = (1.0 + k) + 2.0 * z0[:, None] - 0.5 * y0[None, :] + 0.75 * z0[:, None] * y0[None, :]
        f[1, ..., k] = -f[0, ..., k]
    g = WS.resample_fields(f, L0, M0, L1, M1)
    assert g.shape == (2, L1, M1 + 1, 8)
    z1 = np.arange(L1) / L1
    y1 = np.arange(M1 + 1) / M1
    inside = z1 <= z0[-1]
    for k in range(8):
        want = (1.0 + k) + 2.0 * z1[:, None] - 0.5 * y1[None, :] + 0.75 * z1[:, None] * y1[None, :]
        np.testing.assert_allclose(g[0, inside, :, k], want[inside], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(g[0, ~inside, :, k], np.broadcast_to(g[0, inside, :, k][-1], g[0, ~inside, :, k].shape),
                                   rtol=1e-13)
    np.testing.assert_allclose(g[1], -g[0], rtol=0, atol=0)
    assert np.array_equal(WS.resample_fields(f, L0, M0, L0, M0), f)


def test_fallback_rule_rising_residual_and_budget():
    """A warm case restarts cold when its residual climbs grow-fold above the
    smallest value it reached (a growing mode), or is not finite, or when it
    runs past the step budget; a falling residual keeps running."""
    ws = WS
    least = np.inf
    for c in (3e-4, 5e-6, 2.5e-7, 1.2e-7):
        least, up, over = ws.fallback_rule(c, least, 1000, 10.0, 500000)
        assert not up and not over
    assert least == 1.2e-7
    least, up, _ = ws.fallback_rule(1.1e-6, least, 1000, 10.0, None)
    assert not up and least == 1.2e-7  # 9.2x: still running
    _, up, _ = ws.fallback_rule(1.3e-6, least, 1000, 10.0, None)
    assert up
    _, up, _ = ws.fallback_rule(float("nan"), least, 1000, 10.0, None)
    assert up
    _, up, over = ws.fallback_rule(1e-3, least, 600001, None, 600000)
    assert not up and over


def test_network_file_bf16(tmp_path):
    """save_network_bf16 keeps the weights as the bf16 bit patterns of the
    FP32 masters (round to nearest even, the kernels' rounding of the
    weights' bf16 copies), which decode to FP32 values whose bf16 rounding is
    themselves: the network computes with the same bf16 weights."""
    rng = np.random.default_rng(5)
    w = [rng.standard_normal((4, 9 * 3)).astype(np.float32) for _ in range(2)]
    # Halfway cases of the bf16 rounding, both parities of the kept bit.
    w[0][0, :4] = np.array([0x3F808000, 0x3F818000, 0xBF808001, 0x7F7FFFFF], dtype=np.uint32).view(np.float32)
    b = [rng.standard_normal(4).astype(np.float32) for _ in range(2)]
    src, dst = tmp_path / "net.npz", tmp_path / "net_bf16.npz"
    np.savez(src, w0=w[0], w1=w[1], b0=b[0], b1=b[1], mean=np.arange(8.0), std=np.ones(8),
             train_steps=np.array([10, 20]))
    WS.save_network_bf16(src, dst)
    z = np.load(dst)
    assert sorted(z.files) == ["b0", "b1", "mean", "std", "train_steps", "wb0", "wb1"]
    assert z["wb0"][0, :4].tolist() == [0x3F80, 0x3F82, 0xBF81, 0x7F80]
    for k in range(2):
        assert np.array_equal(z[f"wb{k}"], SG.bf16_bits(w[k]))
        dec = (z[f"wb{k}"].astype(np.uint32) << 16).view(np.float32)
        assert np.array_equal(SG.bf16_bits(dec), z[f"wb{k}"])
        fin = np.isfinite(dec)  # the largest float rounds to +inf, as in the kernels
        assert fin.sum() == w[k].size - (k == 0)
        assert np.all(np.abs(dec - w[k])[fin] <= np.abs(w[k])[fin] * 2.0 ** -8)
        assert np.array_equal(z[f"b{k}"], b[k])
    assert np.array_equal(z["train_steps"], [10, 20])
