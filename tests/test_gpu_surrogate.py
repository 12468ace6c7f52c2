"""GPU tests of the warm-start network's kernels (libmpd_nn.so) against the
float64 oracle (oracle/surrogate_oracle.py) fed the same bf16 operands.

Tolerances: a convolution accumulates exact bf16 x bf16 products in FP32, so
it matches the float64 result to a few FP32 ulps of the accumulated
magnitude (bound written in each test).  Where the kernels round an FP32
value to bf16 (SiLU output, pooled values), an FP32-level difference can move
the rounding by one bf16 ulp (2^-8 relative): those comparisons allow one
ulp.  Whole-network gradients pass through many such roundings and are
compared in norm."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import surrogate_oracle as S  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nn():
    import surrogate
    surrogate.nn_lib()
    return surrogate


def _dev_bf16(a):
    import surrogate
    return surrogate.DeviceArray.from_numpy(a, "bf16")


def _dev_f32(a):
    import surrogate
    return surrogate.DeviceArray.from_numpy(a, "f32")


def _host(t):
    return t.download().astype(np.float64)


def _host_bf16(t):
    return t.download().astype(np.float64)


def _empty(shape, bf16=False):
    import surrogate
    return surrogate.DeviceArray(shape, "bf16" if bf16 else "f32")


# Channel counts that are multiples of 8 take the matrix-core GEMM with the
# 16x16x32 instruction (tile shapes for N = 8, 16, 32, 64, 128, 256 and small
# pixel counts); the others the element-staged kernel.  The weight gradient
# of channel counts that are multiples of 16 takes the LDS row-block kernel
# (wgrad_lds): shapes with several chunks and slices per image, rows of 21
# and of 300 pixels (one row per chunk), and 400-pixel rows, too wide for
# its LDS tiles (the GEMM path then); and rows of 21, 84 and 168 pixels,
# the network's levels.  Gathered channels in 32-channel chunks take the
# three-tap slab kernel on its tiles.
SHAPES = [(2, 8, 5, 3, 5), (3, 12, 10, 8, 32), (1, 16, 8, 96, 32), (2, 6, 5, 40, 70), (1, 5, 3, 256, 130),
          (2, 9, 7, 16, 8), (1, 20, 24, 64, 128), (4, 3, 2, 32, 256), (1, 4, 4, 128, 16), (2, 10, 6, 384, 128),
          (2, 40, 21, 32, 32), (1, 9, 300, 16, 16), (1, 4, 400, 16, 32), (1, 5, 168, 32, 32),
          (1, 6, 84, 64, 64), (2, 13, 21, 64, 48),
          # one-pixel-wide and one-row images (every shifted tap of one axis
          # leaves the image), and 96 channels (the slab kernel's three chunks)
          (8, 40, 1, 32, 32), (3, 1, 9, 64, 32), (1, 6, 10, 96, 32)]


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_passes_match_oracle(nn, shape):
    B, H, W, Ci, Co = shape
    print(f"conv shape B={B} H={H} W={W} Cin={Ci} Cout={Co}", flush=True)
    rng = np.random.default_rng(sum(shape))
    x = S.bf16_round(rng.normal(size=(B, H, W, Ci)))
    w = S.bf16_round(rng.normal(0, 0.3, size=(Co, 3, 3, Ci)))
    b = rng.normal(size=Co).astype(np.float32).astype(np.float64)
    dz = S.bf16_round(rng.normal(size=(B, H, W, Co)))
    L, s = nn.nn_lib(), nn._stream()
    X, Wt, Bt, DZ = _dev_bf16(x), _dev_bf16(w), _dev_f32(b), _dev_bf16(dz)
    Z = _empty((B, H, W, Co))
    nn._ck(L.mpd_nn_conv3x3_fwd(nn._p(X), nn._p(Wt), nn._p(Bt), nn._p(Z), B, H, W, Ci, Co, s))
    DX = _empty((B, H, W, Ci))
    nn._ck(L.mpd_nn_conv3x3_bwd_data(nn._p(DZ), nn._p(Wt), nn._p(DX), B, H, W, Ci, Co, s))
    work = _empty((int(L.mpd_nn_wgrad_work_floats(B, H, W, Ci, Co)),))
    DW = _empty((Co, 3, 3, Ci))
    nn._ck(L.mpd_nn_conv3x3_bwd_weight(nn._p(X), nn._p(DZ), nn._p(DW), nn._p(work), B, H, W, Ci, Co, s))
    # Bound: FP32 accumulation error <= K * 2^-24 * sum |terms| (K terms).
    ref = S.conv_fwd(x, w, b)
    mag = S.conv_fwd(np.abs(x), np.abs(w)) + np.abs(b)
    assert np.all(np.abs(_host(Z) - ref) <= 9 * Ci * 2.0 ** -23 * mag + 1e-30)
    ref = S.conv_bwd_data(dz, w, x.shape)
    mag = S.conv_bwd_data(np.abs(dz), np.abs(w), x.shape)
    assert np.all(np.abs(_host(DX) - ref) <= 9 * Co * 2.0 ** -23 * mag + 1e-30)
    ref = S.conv_bwd_weight(x, dz)
    mag = S.conv_bwd_weight(np.abs(x), np.abs(dz))
    assert np.all(np.abs(_host(DW) - ref) <= B * H * W * 2.0 ** -23 * mag + 1e-30)


def test_elementwise_kernels_match_oracle(nn):
    L, s = nn.nn_lib(), nn._stream()
    rng = np.random.default_rng(7)
    B, H, W, C = 2, 6, 8, 5
    z = rng.normal(0, 2, size=(B, H, W, C)).astype(np.float32).astype(np.float64)
    n = z.size
    Z = _dev_f32(z)
    A = _empty(z.shape, bf16=True)
    nn._ck(L.mpd_nn_silu_fwd(nn._p(Z), nn._p(A), n, s))
    ref = S.silu(z)
    assert np.all(np.abs(_host_bf16(A) - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-30)
    da = rng.normal(size=z.shape).astype(np.float32).astype(np.float64)
    DZB, DZF = _empty(z.shape, bf16=True), _empty(z.shape)
    nn._ck(L.mpd_nn_silu_bwd(nn._p(Z), nn._p(_dev_f32(da)), nn._p(DZB), nn._p(DZF), n, s))
    ref = da * S.silu_grad(z)
    np.testing.assert_allclose(_host(DZF), ref, rtol=2e-6, atol=1e-6)
    assert np.all(np.abs(_host_bf16(DZB) - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-6)

    x = S.bf16_round(rng.normal(size=(B, H, W, C)))
    X = _dev_bf16(x)
    P = _empty((B, H // 2, W // 2, C), bf16=True)
    nn._ck(L.mpd_nn_avgpool2_fwd(nn._p(X), nn._p(P), B, H, W, C, s))
    assert np.array_equal(_host_bf16(P), S.bf16_round(S.pool_fwd(x)))
    dy = rng.normal(size=(B, H // 2, W // 2, C)).astype(np.float32).astype(np.float64)
    DX = _empty((B, H, W, C))
    nn._ck(L.mpd_nn_avgpool2_bwd(nn._p(_dev_f32(dy)), nn._p(DX), B, H, W, C, s))
    assert np.array_equal(_host(DX), S.pool_bwd(dy))
    U = _empty((B, 2 * H, 2 * W, C), bf16=True)
    nn._ck(L.mpd_nn_upsample2_fwd(nn._p(X), nn._p(U), B, H, W, C, s))
    assert np.array_equal(_host_bf16(U), S.up_fwd(x))
    du = rng.normal(size=(B, 2 * H, 2 * W, C)).astype(np.float32).astype(np.float64)
    DU = _empty((B, H, W, C))
    nn._ck(L.mpd_nn_upsample2_bwd(nn._p(_dev_f32(du)), nn._p(DU), B, H, W, C, s))
    np.testing.assert_allclose(_host(DU), S.up_bwd(du), rtol=1e-6, atol=1e-6)

    y2 = S.bf16_round(rng.normal(size=(B, H, W, 3)))
    CT = _empty((B, H, W, C + 3), bf16=True)
    nn._ck(L.mpd_nn_concat_fwd(nn._p(X), C, nn._p(_dev_bf16(y2)), 3, nn._p(CT), B * H * W, s))
    assert np.array_equal(_host_bf16(CT), np.concatenate([x, y2], axis=-1))
    dc = rng.normal(size=(B, H, W, C + 3)).astype(np.float32).astype(np.float64)
    base_b = rng.normal(size=(B, H, W, 3)).astype(np.float32).astype(np.float64)
    DA, DB = _empty((B, H, W, C)), _dev_f32(base_b)
    nn._ck(L.mpd_nn_concat_bwd(nn._p(_dev_f32(dc)), nn._p(DA), C, 0, nn._p(DB), 3, 1, B * H * W, s))
    assert np.array_equal(_host(DA), dc[..., :C])
    np.testing.assert_allclose(_host(DB), base_b + dc[..., C:], rtol=1e-6)

    t = rng.normal(size=z.shape).astype(np.float32).astype(np.float64)
    DY = _empty(z.shape)
    part = _empty((1024,))
    nn._ck(L.mpd_nn_mse(nn._p(Z), nn._p(_dev_f32(t)), nn._p(DY), nn._p(part), n, s))
    loss = _host(part)[:L.mpd_nn_mse_blocks(n)].sum() / n
    rl, rdy = S.mse(z, t)
    assert abs(loss - rl) <= 1e-5 * rl
    np.testing.assert_allclose(_host(DY), rdy, rtol=1e-5, atol=1e-9)

    w0 = rng.normal(size=50).astype(np.float32).astype(np.float64)
    g = rng.normal(size=50).astype(np.float32).astype(np.float64)
    Wd, G, M, V = _dev_f32(w0), _dev_f32(g), _empty((50,)), _empty((50,))
    M.zero()
    V.zero()
    WB = _empty((50,), bf16=True)
    for step in (1, 2):
        nn._ck(L.mpd_nn_adam(nn._p(Wd), nn._p(WB), nn._p(G), nn._p(M), nn._p(V), 50, 1e-2, 0.9, 0.999, 1e-8, step,
                             0.5, s))
    rw, rm, rv = w0, np.zeros(50), np.zeros(50)
    for step in (1, 2):
        rw, rm, rv = S.adam(rw, 0.5 * g, rm, rv, 1e-2, 0.9, 0.999, 1e-8, step)
    np.testing.assert_allclose(_host(Wd), rw, rtol=1e-5, atol=1e-6)
    assert np.array_equal(_host_bf16(WB), S.bf16_round(_host(Wd)))

    # Channel counts of every lane layout: 64 wide with idle lanes (3, 70),
    # C dividing 64 (1, 8, 16, 32), multiples of 64 (128).
    for P, C in ((5, 3), (3000, 70), (40000, 8), (1000, 1), (9000, 16), (20000, 32), (7000, 128)):
        dzb = rng.normal(size=(P, C)).astype(np.float32)
        DB = _empty((C,))
        work = _empty((int(L.mpd_nn_bias_work_floats(P, C)),))
        nn._ck(L.mpd_nn_bias_grad(nn._p(_dev_f32(dzb)), nn._p(DB), nn._p(work), P, C, s))
        ref = dzb.astype(np.float64).sum(axis=0)
        mag = np.abs(dzb.astype(np.float64)).sum(axis=0)
        assert np.all(np.abs(_host(DB) - ref) <= P * 2.0 ** -23 * mag)

    xs = rng.normal(size=70000).astype(np.float32)
    part = _empty((L.mpd_nn_sumsq_blocks(xs.size),))
    nn._ck(L.mpd_nn_sumsq(nn._p(_dev_f32(xs)), nn._p(part), xs.size, s))
    ref = np.sum(xs.astype(np.float64) ** 2)
    assert abs(_host(part).sum() - ref) <= xs.size * 2.0 ** -23 * ref

    src = rng.normal(size=(6, 4)).astype(np.float32)
    idx = np.array([5, 0, 3], np.int32)
    dst = _empty((3, 4))
    I = nn.DeviceArray.from_numpy(idx, "i32")
    nn._ck(L.mpd_nn_gather_rows(nn._p(_dev_f32(src)), nn._p(I), nn._p(dst), 3, 4, s))
    assert np.array_equal(_host(dst), src[idx].astype(np.float64))


@pytest.mark.parametrize("shape", [(2, 9, 7, 16, 8), (1, 20, 24, 64, 128), (2, 6, 5, 40, 70), (1, 5, 168, 32, 32),
                                   (2, 13, 21, 64, 96)])
def test_fused_passes_equal_separate(nn, shape):
    """The fused entry points equal the separate passes bit for bit: the
    forward GEMM with SiLU in its epilogue against the forward then silu_fwd,
    the SiLU backward with the bias gradient (and without activation,
    z = NULL) against silu_bwd / to_bf16 then bias_grad, and the data
    gradient with the SiLU backward below it (bias in another order)."""
    B, H, W, Ci, Co = shape
    print(f"fused passes B={B} H={H} W={W} Cin={Ci} Cout={Co}", flush=True)
    rng = np.random.default_rng(sum(shape) + 1)
    L, s = nn.nn_lib(), nn._stream()
    x = S.bf16_round(rng.normal(size=(B, H, W, Ci)))
    w = S.bf16_round(rng.normal(0, 0.3, size=(Co, 3, 3, Ci)))
    b = rng.normal(size=Co).astype(np.float32).astype(np.float64)
    X, Wt, Bt = _dev_bf16(x), _dev_bf16(w), _dev_f32(b)
    Z1, Z2 = _empty((B, H, W, Co)), _empty((B, H, W, Co))
    A1, A2 = _empty((B, H, W, Co), bf16=True), _empty((B, H, W, Co), bf16=True)
    nn._ck(L.mpd_nn_conv3x3_fwd(nn._p(X), nn._p(Wt), nn._p(Bt), nn._p(Z1), B, H, W, Ci, Co, s))
    nn._ck(L.mpd_nn_silu_fwd(nn._p(Z1), nn._p(A1), Z1.numel(), s))
    nn._ck(L.mpd_nn_conv3x3_fwd_silu(nn._p(X), nn._p(Wt), nn._p(Bt), nn._p(Z2), nn._p(A2), B, H, W, Ci, Co, s))
    assert np.array_equal(Z1.download(), Z2.download())
    assert np.array_equal(A1.download(), A2.download())

    P, n = B * H * W, B * H * W * Co
    da = rng.normal(size=(B, H, W, Co)).astype(np.float32).astype(np.float64)
    DA = _dev_f32(da)
    work = _empty((int(L.mpd_nn_bias_work_floats(P, Co)),))
    for act in (True, False):
        DZB1, DZB2, DZF = _empty((n,), bf16=True), _empty((n,), bf16=True), _empty((n,))
        DB1, DB2 = _empty((Co,)), _empty((Co,))
        if act:
            nn._ck(L.mpd_nn_silu_bwd(nn._p(Z1), nn._p(DA), nn._p(DZB1), nn._p(DZF), n, s))
            nn._ck(L.mpd_nn_bias_grad(nn._p(DZF), nn._p(DB1), nn._p(work), P, Co, s))
        else:
            nn._ck(L.mpd_nn_to_bf16(nn._p(DA), nn._p(DZB1), n, s))
            nn._ck(L.mpd_nn_bias_grad(nn._p(DA), nn._p(DB1), nn._p(work), P, Co, s))
        nn._ck(L.mpd_nn_silu_bwd_bias(nn._p(Z1) if act else None, nn._p(DA), nn._p(DZB2), nn._p(DB2), nn._p(work), P,
                                      Co, s))
        assert np.array_equal(DZB1.download(), DZB2.download()), act
        assert np.array_equal(DB1.download(), DB2.download()), act
        if act:
            ref = (da * S.silu_grad(_host(Z1))).reshape(P, Co)
            mag = np.abs(ref).sum(axis=0)
            assert np.all(np.abs(_host(DB2) - ref.sum(axis=0)) <= P * 2.0 ** -22 * mag + 1e-30)

    # The data gradient with the SiLU backward of the layer below fused: this
    # conv as that layer's consumer (its input's pre-activation z_in), against
    # the data gradient then the SiLU backward with the bias: dz the same
    # bits, the bias gradient within FP32 summation error (another order).
    if Ci % 8 or Co % 8 or (Ci % 96 == 0 and Ci % 128):
        return
    zi = rng.normal(0, 2, size=(B, H, W, Ci)).astype(np.float32).astype(np.float64)
    ZI = _dev_f32(zi)
    dzl = S.bf16_round(rng.normal(size=(B, H, W, Co)))
    DZL = _dev_bf16(dzl)
    DX = _empty((B, H, W, Ci))
    nn._ck(L.mpd_nn_conv3x3_bwd_data(nn._p(DZL), nn._p(Wt), nn._p(DX), B, H, W, Ci, Co, s))
    n_in = B * H * W * Ci
    DZ1, DZ2 = _empty((n_in,), bf16=True), _empty((n_in,), bf16=True)
    DB1, DB2 = _empty((Ci,)), _empty((Ci,))
    work_b = _empty((int(L.mpd_nn_bias_work_floats(P, Ci)),))
    nn._ck(L.mpd_nn_silu_bwd_bias(nn._p(ZI), nn._p(DX), nn._p(DZ1), nn._p(DB1), nn._p(work_b), P, Ci, s))
    work_f = _empty((int(L.mpd_nn_bwd_data_silu_work_floats(B, H, W, Ci, Co)),))
    nn._ck(L.mpd_nn_conv3x3_bwd_data_silu(nn._p(DZL), nn._p(Wt), nn._p(ZI), nn._p(DZ2), nn._p(DB2), nn._p(work_f),
                                          B, H, W, Ci, Co, s))
    assert np.array_equal(DZ1.download(), DZ2.download())
    d = (_host(DX) * S.silu_grad(zi)).reshape(P, Ci)
    mag = np.abs(d).sum(axis=0)
    assert np.all(np.abs(_host(DB2) - _host(DB1)) <= 2 * P * 2.0 ** -23 * mag + 1e-30)


@pytest.mark.parametrize("B,H,W,Cu,Cs", [(2, 3, 5, 16, 8), (1, 25, 21, 64, 32), (2, 4, 6, 8, 24), (1, 50, 84, 64, 32)])
def test_fused_joins_equal_separate(nn, B, H, W, Cu, Cs):
    """The decoder's skip joins in one pass equal the separate passes bit for
    bit: upsample + concatenation forward, their backward, and the encoder's
    pooling backward plus the skip gradient (H, W the coarse sizes)."""
    print(f"fused joins B={B} H={H} W={W} Cu={Cu} Cs={Cs}", flush=True)
    rng = np.random.default_rng(B * H * W + Cu + Cs)
    L, s = nn.nn_lib(), nn._stream()
    Hf, Wf, P = 2 * H, 2 * W, B * 4 * H * W
    x = S.bf16_round(rng.normal(size=(B, H, W, Cu)))
    sk = S.bf16_round(rng.normal(size=(B, Hf, Wf, Cs)))
    X, SK = _dev_bf16(x), _dev_bf16(sk)
    U, Y1, Y2 = (_empty((B, Hf, Wf, Cu), bf16=True), _empty((B, Hf, Wf, Cu + Cs), bf16=True),
                 _empty((B, Hf, Wf, Cu + Cs), bf16=True))
    nn._ck(L.mpd_nn_upsample2_fwd(nn._p(X), nn._p(U), B, H, W, Cu, s))
    nn._ck(L.mpd_nn_concat_fwd(nn._p(U), Cu, nn._p(SK), Cs, nn._p(Y1), P, s))
    nn._ck(L.mpd_nn_upcat_fwd(nn._p(X), Cu, nn._p(SK), Cs, nn._p(Y2), B, H, W, s))
    assert np.array_equal(Y1.download(), Y2.download())
    assert np.array_equal(_host_bf16(Y2), np.concatenate([S.up_fwd(x), sk], axis=-1))

    dy = rng.normal(size=(B, Hf, Wf, Cu + Cs)).astype(np.float32)
    DY = _dev_f32(dy)
    T, DX1, DS1 = _empty((B, Hf, Wf, Cu)), _empty((B, H, W, Cu)), _empty((B, Hf, Wf, Cs))
    DX2, DS2 = _empty((B, H, W, Cu)), _empty((B, Hf, Wf, Cs))
    nn._ck(L.mpd_nn_concat_bwd(nn._p(DY), nn._p(T), Cu, 0, nn._p(DS1), Cs, 0, P, s))
    nn._ck(L.mpd_nn_upsample2_bwd(nn._p(T), nn._p(DX1), B, H, W, Cu, s))
    nn._ck(L.mpd_nn_upcat_bwd(nn._p(DY), nn._p(DX2), Cu, nn._p(DS2), Cs, B, H, W, s))
    assert np.array_equal(DX1.download(), DX2.download())
    assert np.array_equal(DS1.download(), DS2.download())
    np.testing.assert_allclose(_host(DX2), S.up_bwd(dy[..., :Cu].astype(np.float64)), rtol=1e-6, atol=1e-6)

    g = rng.normal(size=(B, H, W, Cs)).astype(np.float32)
    G = _dev_f32(g)
    PB1, PB2 = _empty((B, Hf, Wf, Cs)), _empty((B, Hf, Wf, Cs))
    nn._ck(L.mpd_nn_avgpool2_bwd(nn._p(G), nn._p(PB1), B, Hf, Wf, Cs, s))
    nn._ck(L.mpd_nn_add(nn._p(PB1), nn._p(DS2), B * Hf * Wf * Cs, s))
    nn._ck(L.mpd_nn_pool_bwd_add(nn._p(G), nn._p(DS2), nn._p(PB2), B, Hf, Wf, Cs, s))
    assert np.array_equal(PB1.download(), PB2.download())
    np.testing.assert_allclose(_host(PB2), S.pool_bwd(g.astype(np.float64)) + dy[..., Cu:], rtol=1e-6, atol=1e-7)

    # ... and with the SiLU backward and bias slices in the same pass.
    z = rng.normal(size=(B, Hf, Wf, Cs)).astype(np.float32)
    Z = _dev_f32(z)
    DZ1, DZ2 = _empty((P * Cs,), bf16=True), _empty((P * Cs,), bf16=True)
    DB1, DB2 = _empty((Cs,)), _empty((Cs,))
    work = _empty((int(L.mpd_nn_bias_work_floats(P, Cs)),))
    nn._ck(L.mpd_nn_silu_bwd_bias(nn._p(Z), nn._p(PB2), nn._p(DZ1), nn._p(DB1), nn._p(work), P, Cs, s))
    nn._ck(L.mpd_nn_pool_silu_bwd_bias(nn._p(G), nn._p(DS2), nn._p(Z), nn._p(DZ2), nn._p(DB2), nn._p(work), B, Hf,
                                       Wf, Cs, s))
    assert np.array_equal(DZ1.download(), DZ2.download())
    assert np.array_equal(DB1.download(), DB2.download())
    d = (_host(PB2) * S.silu_grad(z)).reshape(P, Cs)
    assert np.all(np.abs(_host(DB2) - d.sum(axis=0)) <= 4 * P * 2.0 ** -23 * np.abs(d).sum(axis=0) + 1e-30)


def test_adam_multi_equals_separate(nn):
    """One multi-tensor Adam launch (35 tensors: two launches of the table,
    sizes 0 to 5000, some without a bf16 copy) equals mpd_nn_adam on each
    tensor bit for bit, over three steps."""
    import ctypes as C
    L, s = nn.nn_lib(), nn._stream()
    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(0, 5000, size=35)]
    sizes[3] = 0
    sizes[7] = 1
    sets = []
    for _ in range(2):
        ts = []
        for k, n in enumerate(sizes):
            w0 = rng.normal(size=max(n, 1)).astype(np.float32)[:n] if n else np.zeros(0, np.float32)
            ts.append(dict(w=_dev_f32(w0) if n else _empty((1,)), wb=_empty((max(n, 1),), bf16=True) if k % 3 else None,
                           m=_empty((max(n, 1),)), v=_empty((max(n, 1),))))
            ts[-1]["m"].zero()
            ts[-1]["v"].zero()
        sets.append(ts)
    # Same initial weights in both sets.
    for a, b in zip(*sets):
        if a["w"].numel() and b["w"].numel():
            b["w"].upload(a["w"].download())
    grads = [_dev_f32(rng.normal(size=max(n, 1)).astype(np.float32)) for n in sizes]
    ptrs = lambda key, ts: (C.c_void_p * len(ts))(*[(t[key].ptr if t[key] is not None else None) for t in ts])
    nn_arr = (C.c_longlong * len(sizes))(*sizes)
    gp = (C.c_void_p * len(sizes))(*[g.ptr for g in grads])
    for step in (1, 2, 3):
        for k, n in enumerate(sizes):
            t = sets[0][k]
            nn._ck(L.mpd_nn_adam(nn._p(t["w"]), nn._p(t["wb"]) if t["wb"] is not None else None, nn._p(grads[k]),
                                 nn._p(t["m"]), nn._p(t["v"]), n, 1e-3, 0.9, 0.999, 1e-8, step, 0.7, s))
        nn._ck(L.mpd_nn_adam_multi(len(sizes), ptrs("w", sets[1]), ptrs("wb", sets[1]), gp, ptrs("m", sets[1]),
                                   ptrs("v", sets[1]), nn_arr, 1e-3, 0.9, 0.999, 1e-8, step, 0.7, s))
    for k, n in enumerate(sizes):
        if n == 0:
            continue
        for key in ("w", "m", "v", "wb"):
            a, b = sets[0][k][key], sets[1][k][key]
            if a is None:
                continue
            assert np.array_equal(a.download()[:n], b.download()[:n]), (k, key)


@pytest.mark.parametrize("base,H,W,B", [(8, 16, 8, 2), (32, 16, 8, 2), (32, 24, 40, 1), (24, 16, 8, 1)])
def test_unet_forward_backward_match_oracle(nn, base, H, W, B):
    """The whole encoder-decoder (4 levels) on the GPU against the oracle with
    the same bf16 roundings: the loss, and every parameter gradient (in
    norm: 1e-2 relative, cosine > 0.9999).  Base 24 puts 96 channels where
    the layers chain directly (the fused data gradient's fallback)."""
    print(f"unet base={base} grid={H}x{W} batch={B}", flush=True)
    rng = np.random.default_rng(base + H)
    net = nn.UNet(H, W, base=base, seed=3)
    ws = [rng.normal(0, np.sqrt(1.0 / (9 * ci)), size=(co, 3, 3, ci)) for ci, co in net.layers]
    bs = [rng.normal(0, 0.05, size=co) for _, co in net.layers]
    net.load_params(ws, bs)
    x = S.bf16_round(rng.normal(size=(B, H, W, 8)))
    t = rng.normal(size=(B, H, W, 8)).astype(np.float32).astype(np.float64)
    y = net.forward(_dev_bf16(x), B)
    T = _dev_f32(t)
    DY = _empty((B, H, W, 8))
    loss = net.mse(y, T, DY)
    net.backward(DY, B)
    params = [(S.bf16_round(np.asarray(w, np.float32)), np.asarray(b, np.float32).astype(np.float64))
              for w, b in zip(ws, bs)]
    ry, tape = S.unet_forward(x, params, bf16=True)
    rl, rdy = S.mse(ry, t)
    assert abs(loss - rl) <= 1e-3 * rl, (loss, rl)
    yg = _host(y)
    assert np.linalg.norm(yg - ry) <= 2e-3 * np.linalg.norm(ry)
    grads = S.unet_backward(rdy, tape, bf16=True)
    for k, (gw, gb) in enumerate(grads):
        for got, ref in ((_host(net.gw[k]), gw), (_host(net.gb[k]), gb)):
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            cos = np.sum(got * ref) / (np.linalg.norm(got) * np.linalg.norm(ref))
            assert rel < 1e-2 and cos > 0.9999, (k, rel, cos)


def test_training_fits_a_smooth_field(nn):
    """Adam on the GPU kernels fits a smooth parameter-dependent field: the
    loss falls by more than 20x in 150 steps."""
    H, W, B = 16, 8, 16
    rng = np.random.default_rng(11)
    p = rng.uniform(-1, 1, size=(B, 3))
    yy, xx = np.meshgrid(np.linspace(-1, 1, H), np.linspace(-1, 1, W), indexing="ij")
    x = np.zeros((B, H, W, 8))
    x[..., :3] = p[:, None, None, :]
    x[..., 3] = yy
    x[..., 4] = xx
    t = np.zeros((B, H, W, 8))
    for c in range(8):
        t[..., c] = np.sin((c + 1) * 0.5 * yy + p[:, 0, None, None]) * (1 + 0.5 * p[:, 1, None, None] * xx)
    net = nn.UNet(H, W, base=16, seed=1)
    X, T = _dev_bf16(x), _dev_f32(t)
    DY = _empty((B, H, W, 8))
    losses = []
    for it in range(150):
        y = net.forward(X, B)
        losses.append(net.mse(y, T, DY))
        net.backward(DY, B)
        net.adam_step(2e-3)
    print(f"loss {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < losses[0] / 20
    assert np.all(np.isfinite(losses))
