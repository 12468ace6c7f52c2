"""Learned warm start for steady-state sweeps of the LF solver.

A BF16 convolutional encoder-decoder (U-Net: `levels` resolution levels,
`base` channels at the finest, doubling per level, 3x3 convolutions, SiLU,
2x2 average pooling, nearest x2 upsampling with skip concatenation, FP32
accumulation) maps the sweep parameters of a case -- inflow density,
inflow velocity, B_phi amplitude and the two wall-shape amplitudes of the
converging-diverging channel -- to the 8 steady-state physical fields on the
solver's grid.  Every forward and backward operation runs on the hand-written
gfx950 kernels of libmpd_nn.so (include/mpd_nn.h), on device memory that
library allocates; the host side is numpy.  The prediction becomes the initial state of the
plain FP64 LF loop, which relaxes it to the convergence threshold: the solver
decides, and a case whose residual stalls falls back to a cold start.

No counterpart in the reference: its runs start from the expression or VTK
initial condition (src/main.cpp:117-148).
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NN_LIB_PATH = os.path.join(HERE, "libmpd_nn.so")
if HERE not in sys.path:
    sys.path.insert(0, HERE)

from sweep import PARAM_HI, PARAM_LO  # noqa: E402  (the sweep's parameter box)

# Input features per pixel: the 5 parameters (scaled to [-1, 1]), z, the
# radial grid fraction and the radius.
N_FEATURES = 8
# Output channels: the physical fields the warm start sets (Fields member
# names), in this order.
OUT_FIELDS = ["rho", "v_z", "v_r", "v_phi", "e", "H_z", "H_r", "H_phi"]
PHYS_INDEX = {"rho": 0, "v_z": 1, "v_r": 2, "v_phi": 3, "e": 4, "p": 5, "P": 6, "H_z": 7, "H_r": 8, "H_phi": 9}

_nn = None
_vp = C.c_void_p


def nn_lib():
    """Load libmpd_nn.so (raises if it was not built: no CPU fallback)."""
    global _nn
    if _nn is not None:
        return _nn
    if not os.path.exists(NN_LIB_PATH):
        raise RuntimeError(f"{NN_LIB_PATH} is missing: build it first")
    L = C.CDLL(NN_LIB_PATH)
    i, ll, f = C.c_int, C.c_longlong, C.c_float
    sig = {
        "mpd_nn_conv3x3_fwd": [_vp, _vp, _vp, _vp, i, i, i, i, i, _vp],
        "mpd_nn_conv3x3_fwd_silu": [_vp, _vp, _vp, _vp, _vp, i, i, i, i, i, _vp],
        "mpd_nn_conv3x3_bwd_data": [_vp, _vp, _vp, i, i, i, i, i, _vp],
        "mpd_nn_conv3x3_bwd_data_silu": [_vp, _vp, _vp, _vp, _vp, _vp, i, i, i, i, i, _vp],
        "mpd_nn_bwd_data_silu_work_floats": [i, i, i, i, i],
        "mpd_nn_conv3x3_bwd_weight": [_vp, _vp, _vp, _vp, i, i, i, i, i, _vp],
        "mpd_nn_wgrad_work_floats": [i, i, i, i, i],
        "mpd_nn_bias_grad": [_vp, _vp, _vp, i, i, _vp],
        "mpd_nn_bias_work_floats": [i, i],
        "mpd_nn_silu_fwd": [_vp, _vp, ll, _vp],
        "mpd_nn_silu_bwd": [_vp, _vp, _vp, _vp, ll, _vp],
        "mpd_nn_silu_bwd_bias": [_vp, _vp, _vp, _vp, _vp, i, i, _vp],
        "mpd_nn_to_bf16": [_vp, _vp, ll, _vp],
        "mpd_nn_avgpool2_fwd": [_vp, _vp, i, i, i, i, _vp],
        "mpd_nn_avgpool2_bwd": [_vp, _vp, i, i, i, i, _vp],
        "mpd_nn_upsample2_fwd": [_vp, _vp, i, i, i, i, _vp],
        "mpd_nn_upsample2_bwd": [_vp, _vp, i, i, i, i, _vp],
        "mpd_nn_concat_fwd": [_vp, i, _vp, i, _vp, ll, _vp],
        "mpd_nn_concat_bwd": [_vp, _vp, i, i, _vp, i, i, ll, _vp],
        "mpd_nn_add": [_vp, _vp, ll, _vp],
        "mpd_nn_upcat_fwd": [_vp, i, _vp, i, _vp, i, i, i, _vp],
        "mpd_nn_upcat_bwd": [_vp, _vp, i, _vp, i, i, i, i, _vp],
        "mpd_nn_pool_bwd_add": [_vp, _vp, _vp, i, i, i, i, _vp],
        "mpd_nn_pool_silu_bwd_bias": [_vp, _vp, _vp, _vp, _vp, _vp, i, i, i, i, _vp],
        "mpd_nn_mse_blocks": [ll],
        "mpd_nn_mse": [_vp, _vp, _vp, _vp, ll, _vp],
        "mpd_nn_adam": [_vp, _vp, _vp, _vp, _vp, ll, f, f, f, f, i, f, _vp],
        "mpd_nn_adam_multi": [i, _vp, _vp, _vp, _vp, _vp, _vp, f, f, f, f, i, f, _vp],
        "mpd_nn_sumsq_blocks": [ll],
        "mpd_nn_sumsq": [_vp, _vp, ll, _vp],
        "mpd_nn_gather_rows": [_vp, _vp, _vp, i, ll, _vp],
        "mpd_nn_malloc": [C.POINTER(_vp), ll],
        "mpd_nn_free": [_vp],
        "mpd_nn_memcpy_h2d": [_vp, _vp, ll],
        "mpd_nn_memcpy_d2h": [_vp, _vp, ll],
        "mpd_nn_memset": [_vp, i, ll],
        "mpd_nn_synchronize": [],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_longlong if name.endswith("_work_floats") else C.c_int
    L.mpd_nn_last_error.restype = C.c_char_p
    _nn = L
    return L


# Every symbol include/mpd_nn.h declares.
NN_EXPORTS = ["mpd_nn_last_error", "mpd_nn_conv3x3_fwd", "mpd_nn_conv3x3_fwd_silu", "mpd_nn_conv3x3_bwd_data",
              "mpd_nn_conv3x3_bwd_data_silu", "mpd_nn_bwd_data_silu_work_floats",
              "mpd_nn_wgrad_work_floats", "mpd_nn_silu_bwd_bias", "mpd_nn_upcat_fwd", "mpd_nn_upcat_bwd",
              "mpd_nn_pool_bwd_add", "mpd_nn_pool_silu_bwd_bias",
              "mpd_nn_conv3x3_bwd_weight", "mpd_nn_bias_work_floats", "mpd_nn_bias_grad", "mpd_nn_silu_fwd", "mpd_nn_silu_bwd",
              "mpd_nn_to_bf16", "mpd_nn_avgpool2_fwd", "mpd_nn_avgpool2_bwd", "mpd_nn_upsample2_fwd",
              "mpd_nn_upsample2_bwd", "mpd_nn_concat_fwd", "mpd_nn_concat_bwd", "mpd_nn_add", "mpd_nn_mse_blocks",
              "mpd_nn_mse", "mpd_nn_adam", "mpd_nn_adam_multi", "mpd_nn_sumsq_blocks", "mpd_nn_sumsq", "mpd_nn_gather_rows", "mpd_nn_malloc", "mpd_nn_free",
              "mpd_nn_memcpy_h2d", "mpd_nn_memcpy_d2h", "mpd_nn_memset", "mpd_nn_synchronize"]


def _ck(rc):
    if rc != 0:
        raise RuntimeError(f"libmpd_nn: [{rc}] {nn_lib().mpd_nn_last_error().decode(errors='replace')}")


def _p(t):
    return C.c_void_p(t.ptr)


def _stream():
    # Every kernel of this module runs on the null stream, in call order.
    return None


def bf16_bits(x):
    """bfloat16 bit patterns (uint16) of float values, rounded to nearest
    even (the rounding of the kernels' FP32 -> bf16 conversion)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bf16_to_float(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class DeviceArray:
    """A device buffer allocated by libmpd_nn.so: 'f32' (float32), 'bf16'
    (bfloat16 bit patterns) or 'i32' (int32), C-contiguous."""

    _ITEM = {"f32": 4, "bf16": 2, "i32": 4}

    def __init__(self, shape, kind="f32"):
        if isinstance(shape, (int, np.integer)):
            shape = (shape,)
        self.shape = tuple(int(v) for v in shape)
        self.kind = kind
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self._ITEM[kind]
        p = C.c_void_p()
        _ck(nn_lib().mpd_nn_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                nn_lib().mpd_nn_free(C.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None

    def numel(self):
        return self.size

    @classmethod
    def from_numpy(cls, a, kind="f32"):
        d = cls(np.shape(a), kind)
        d.upload(a)
        return d

    def upload(self, a):
        if self.kind == "f32":
            h = np.ascontiguousarray(a, dtype=np.float32)
        elif self.kind == "bf16":
            h = bf16_bits(a)
        else:
            h = np.ascontiguousarray(a, dtype=np.int32)
        if h.size != self.size:
            raise ValueError(f"upload of {h.size} values into a buffer of {self.size}")
        _ck(nn_lib().mpd_nn_memcpy_h2d(C.c_void_p(self.ptr), h.ctypes.data_as(C.c_void_p), self.nbytes))

    def download(self):
        """Host copy (float32 values for 'f32' and 'bf16', int32 for 'i32')."""
        dt = {"f32": np.float32, "bf16": np.uint16, "i32": np.int32}[self.kind]
        h = np.empty(self.shape, dtype=dt)
        _ck(nn_lib().mpd_nn_memcpy_d2h(h.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes))
        return bf16_to_float(h) if self.kind == "bf16" else h

    def zero(self):
        _ck(nn_lib().mpd_nn_memset(C.c_void_p(self.ptr), 0, self.nbytes))


def unet_layers(cin, cout, base=32, levels=4):
    """(Cin, Cout) of every conv in execution order (encoder levels: two
    convs each; decoder levels: two each, the first after the skip
    concatenation; then the output head)."""
    ch = [base * 2 ** i for i in range(levels)]
    layers, c = [], cin
    for i in range(levels):
        layers += [(c, ch[i]), (ch[i], ch[i])]
        c = ch[i]
    for i in range(levels - 2, -1, -1):
        layers += [(c + ch[i], ch[i]), (ch[i], ch[i])]
        c = ch[i]
    layers.append((c, cout))
    return layers


def padded(n, levels=4):
    """n rounded up to a multiple of 2^(levels-1): the size the network's
    pooling levels need."""
    q = 2 ** (levels - 1)
    return (n + q - 1) // q * q


def pad_grid(a, H, W):
    """(N, h, w, C) array edge-replicated to (N, H, W, C) (H >= h, W >= w)."""
    a = np.asarray(a)
    h, w = a.shape[1], a.shape[2]
    if (h, w) == (H, W):
        return a
    return np.pad(a, ((0, 0), (0, H - h), (0, W - w), (0, 0)), mode="edge")


class UNet:
    """The encoder-decoder on the GPU: FP32 master weights with bf16 copies,
    Adam moments, and the activations of the last forward pass (kept for the
    backward pass).  A grid of h x w cells runs on the network's padded size
    H x W (each rounded up to a multiple of 2^(levels-1)): inputs and
    training targets are edge-replicated to it (pad_grid) and outputs cropped
    back to h x w."""

    def __init__(self, h, w, cin=N_FEATURES, cout=len(OUT_FIELDS), base=32, levels=4, seed=0):
        self.h, self.w = h, w
        H, W = padded(h, levels), padded(w, levels)
        self.H, self.W, self.cin, self.cout, self.base, self.levels = H, W, cin, cout, base, levels
        self.layers = unet_layers(cin, cout, base, levels)
        rng = np.random.default_rng(seed)
        self.w, self.b, self.wb, self.gw, self.gb = [], [], [], [], []
        self.mw, self.vw, self.mb, self.vb = [], [], [], []
        for k, (ci, co) in enumerate(self.layers):
            std = math.sqrt(2.0 / (9 * ci)) * (0.1 if k == len(self.layers) - 1 else 1.0)
            self.w.append(DeviceArray.from_numpy(rng.normal(0.0, std, size=(co, 3, 3, ci))))
            self.b.append(DeviceArray.from_numpy(np.zeros(co)))
            self.wb.append(DeviceArray((co, 3, 3, ci), "bf16"))
            for lst, shape in ((self.gw, (co, 3, 3, ci)), (self.gb, (co,)), (self.mw, (co, 3, 3, ci)),
                               (self.vw, (co, 3, 3, ci)), (self.mb, (co,)), (self.vb, (co,))):
                d = DeviceArray(shape)
                d.zero()
                lst.append(d)
        self.step = 0
        self._sync_bf16()
        self._B = None

    # -- parameters ------------------------------------------------------
    def _sync_bf16(self):
        L = nn_lib()
        for w, wb in zip(self.w, self.wb):
            _ck(L.mpd_nn_to_bf16(_p(w), _p(wb), w.numel(), _stream()))

    def state_dict(self):
        return {"layers": self.layers, "w": [w.download() for w in self.w], "b": [b.download() for b in self.b]}

    def load_params(self, ws, bs):
        for k in range(len(self.layers)):
            self.w[k].upload(ws[k])
            self.b[k].upload(bs[k])
        self._sync_bf16()

    def snapshot(self):
        """Host copies of the FP32 weights and biases."""
        return [w.download() for w in self.w], [b.download() for b in self.b]

    def restore(self, state):
        """Weights and biases from a snapshot; Adam starts afresh (zero moments,
        step count 0)."""
        self.load_params(*state)
        for t in self.mw + self.vw + self.mb + self.vb:
            t.zero()
        self.step = 0

    def n_params(self):
        return sum(w.numel() + b.numel() for w, b in zip(self.w, self.b))

    # -- buffers -----------------------------------------------------------
    def _alloc(self, B):
        if self._B == B:
            return
        H, W = self.H, self.W
        self.lay = []  # per conv: dict(h, w, ci, co, a_in, z, a_out, act)
        h, w = H, W
        ch = [self.base * 2 ** i for i in range(self.levels)]
        k = 0
        self.x_in = DeviceArray((B, H, W, self.cin), "bf16")
        self.pools, self.cats = [], []
        for i in range(self.levels):
            for j in range(2):
                ci, co = self.layers[k]
                self.lay.append(dict(h=h, w=w, ci=ci, co=co, act=True,
                                     z=DeviceArray((B, h, w, co)),
                                     a_out=DeviceArray((B, h, w, co), "bf16")))
                k += 1
            if i < self.levels - 1:
                self.pools.append(DeviceArray((B, h // 2, w // 2, ch[i]), "bf16"))
                h, w = h // 2, w // 2
        for i in range(self.levels - 2, -1, -1):
            h, w = h * 2, w * 2
            cu = self.layers[k][0] - ch[i]
            self.cats.append(DeviceArray((B, h, w, cu + ch[i]), "bf16"))
            for j in range(2):
                ci, co = self.layers[k]
                self.lay.append(dict(h=h, w=w, ci=ci, co=co, act=True,
                                     z=DeviceArray((B, h, w, co)),
                                     a_out=DeviceArray((B, h, w, co), "bf16")))
                k += 1
        ci, co = self.layers[k]
        self.lay.append(dict(h=h, w=w, ci=ci, co=co, act=False,
                             z=DeviceArray((B, h, w, co)), a_out=None))
        L = nn_lib()
        work = max(L.mpd_nn_wgrad_work_floats(B, d["h"], d["w"], d["ci"], d["co"]) for d in self.lay)
        self.work = DeviceArray(int(work))
        bwork = max(L.mpd_nn_bias_work_floats(B * d["h"] * d["w"], d["co"]) for d in self.lay)
        self.bwork = DeviceArray(int(bwork))
        big = max(B * d["h"] * d["w"] * max(d["ci"], d["co"]) for d in self.lay)
        # Gradient scratch: two ping-pong FP32 buffers, two bf16 dz buffers
        # and the fused data gradient's bias slices, and per-skip gradient
        # buffers.
        self.g0 = DeviceArray(big)
        self.g1 = DeviceArray(big)
        self._dzb = [DeviceArray(big, "bf16"), DeviceArray(big, "bf16")]  # ping-pong: fused data gradients
        self._dzi = 0
        dwork = max(L.mpd_nn_bwd_data_silu_work_floats(B, d["h"], d["w"], d["ci"], d["co"]) for d in self.lay)
        self.dwork = DeviceArray(int(dwork))
        self.dskip = [DeviceArray(B * H // 2 ** i * W // 2 ** i * ch[i])
                      for i in range(self.levels - 1)]
        self.loss_part = DeviceArray(1024)
        self._B = B

    # -- forward -------------------------------------------------------------
    def forward(self, x_bf16, B):
        """x_bf16: (B, H, W, cin) bf16 tensor.  Returns the FP32 output
        (B, H, W, cout) (a view of an internal buffer)."""
        self._alloc(B)
        L, s = nn_lib(), _stream()
        a = x_bf16
        k = 0
        skips = []
        for i in range(self.levels):
            for j in range(2):
                d = self.lay[k]
                d["a_in"] = a
                _ck(L.mpd_nn_conv3x3_fwd_silu(_p(a), _p(self.wb[k]), _p(self.b[k]), _p(d["z"]), _p(d["a_out"]), B,
                                              d["h"], d["w"], d["ci"], d["co"], s))
                a = d["a_out"]
                k += 1
            if i < self.levels - 1:
                skips.append(a)
                d = self.lay[k - 1]
                _ck(L.mpd_nn_avgpool2_fwd(_p(a), _p(self.pools[i]), B, d["h"], d["w"], d["co"], s))
                a = self.pools[i]
        for n, i in enumerate(range(self.levels - 2, -1, -1)):
            d = self.lay[k]
            hh, ww = d["h"] // 2, d["w"] // 2
            cu = d["ci"] - skips[i].shape[-1]
            _ck(L.mpd_nn_upcat_fwd(_p(a), cu, _p(skips[i]), skips[i].shape[-1], _p(self.cats[n]), B, hh, ww, s))
            a = self.cats[n]
            for j in range(2):
                d = self.lay[k]
                d["a_in"] = a
                _ck(L.mpd_nn_conv3x3_fwd_silu(_p(a), _p(self.wb[k]), _p(self.b[k]), _p(d["z"]), _p(d["a_out"]), B,
                                              d["h"], d["w"], d["ci"], d["co"], s))
                a = d["a_out"]
                k += 1
        d = self.lay[k]
        d["a_in"] = a
        _ck(L.mpd_nn_conv3x3_fwd(_p(a), _p(self.wb[k]), _p(self.b[k]), _p(d["z"]), B, d["h"], d["w"], d["ci"],
                                 d["co"], s))
        return d["z"]

    # -- backward ------------------------------------------------------------
    def _conv_back(self, k, da, dx_out, B, fuse=False):
        """Gradients of conv k.  da = dL/d(output) (FP32 tensor), or None when
        this layer's dz and bias gradient were already written by the layer
        above's data gradient (fused).  dx_out: dL/d(input) (FP32), or None
        for the first conv; with fuse the data gradient instead writes conv
        k - 1's dz (bf16) and bias gradient directly (its output feeds this
        conv's input, with nothing in between)."""
        L, s = nn_lib(), _stream()
        d = self.lay[k]
        dzb = self._dzb[self._dzi]
        if da is not None:
            z = _p(d["z"]) if d["act"] else None  # no activation: dz = da
            _ck(L.mpd_nn_silu_bwd_bias(z, _p(da), _p(dzb), _p(self.gb[k]), _p(self.bwork), B * d["h"] * d["w"],
                                       d["co"], s))
        _ck(L.mpd_nn_conv3x3_bwd_weight(_p(d["a_in"]), _p(dzb), _p(self.gw[k]), _p(self.work), B, d["h"], d["w"],
                                        d["ci"], d["co"], s))
        if fuse:
            below = self.lay[k - 1]
            assert below["act"] and below["co"] == d["ci"]
            self._dzi ^= 1
            ci, co = d["ci"], d["co"]
            if ci % 8 == 0 and co % 8 == 0 and not (ci % 96 == 0 and ci % 128):
                _ck(L.mpd_nn_conv3x3_bwd_data_silu(_p(dzb), _p(self.wb[k]), _p(below["z"]),
                                                   _p(self._dzb[self._dzi]), _p(self.gb[k - 1]), _p(self.dwork), B,
                                                   d["h"], d["w"], ci, co, s))
            else:
                # Channel counts the fused entry does not take: the same
                # results through dx and the separate SiLU backward.
                if getattr(self, "_gtmp", None) is None or self._gtmp.numel() < self.g0.numel():
                    self._gtmp = DeviceArray(self.g0.numel())
                _ck(L.mpd_nn_conv3x3_bwd_data(_p(dzb), _p(self.wb[k]), _p(self._gtmp), B, d["h"], d["w"], ci, co,
                                              s))
                _ck(L.mpd_nn_silu_bwd_bias(_p(below["z"]), _p(self._gtmp), _p(self._dzb[self._dzi]),
                                           _p(self.gb[k - 1]), _p(self.bwork), B * below["h"] * below["w"],
                                           below["co"], s))
        elif dx_out is not None:
            _ck(L.mpd_nn_conv3x3_bwd_data(_p(dzb), _p(self.wb[k]), _p(dx_out), B, d["h"], d["w"], d["ci"],
                                          d["co"], s))

    def backward(self, dy, B):
        """dy: dL/d(output) FP32 (B, H, W, cout).  Fills the gradients.  Where a
        conv's input is the previous conv's output (the head and each level's
        second conv), its data gradient also forms the previous conv's dz."""
        L, s = nn_lib(), _stream()
        k = len(self.lay) - 1
        cur, nxt = self.g0, self.g1
        self._dzi = 0
        self._conv_back(k, dy, None, B, fuse=True)  # head -> finest decoder's second conv
        k -= 1
        ch = [self.base * 2 ** i for i in range(self.levels)]
        for i in range(self.levels - 1):  # decoder levels, finest first
            self._conv_back(k, None if i == 0 else cur, None, B, fuse=True)
            k -= 1
            d = self.lay[k]
            self._conv_back(k, None, nxt, B)
            cur, nxt = nxt, cur
            k -= 1
            cu = d["ci"] - ch[i]
            _ck(L.mpd_nn_upcat_bwd(_p(cur), _p(nxt), cu, _p(self.dskip[i]), ch[i], B, d["h"] // 2, d["w"] // 2, s))
            cur, nxt = nxt, cur
        # bottleneck
        self._conv_back(k, cur, None, B, fuse=True)
        k -= 1
        self._conv_back(k, None, nxt, B)
        cur, nxt = nxt, cur
        k -= 1
        for i in range(self.levels - 2, -1, -1):  # encoder levels, coarsest first
            d = self.lay[k]
            assert d["act"]
            # Pooling backward + skip gradient + SiLU backward + bias slices
            # in one pass: this conv's dz, with no FP32 gradient in between.
            _ck(L.mpd_nn_pool_silu_bwd_bias(_p(cur), _p(self.dskip[i]), _p(d["z"]), _p(self._dzb[self._dzi]),
                                            _p(self.gb[k]), _p(self.bwork), B, d["h"], d["w"], d["co"], s))
            self._conv_back(k, None, None, B, fuse=True)
            k -= 1
            self._conv_back(k, None, nxt if k > 0 else None, B)
            cur, nxt = nxt, cur
            k -= 1
        assert k == -1

    def grad_norm(self):
        """Global L2 norm of all parameter gradients (partial sums on the GPU,
        added on the host in a fixed order)."""
        L, s = nn_lib(), _stream()
        if getattr(self, "_gparts", None) is None:
            self._gslots = []
            off = 0
            for t in self.gw + self.gb:
                nb = L.mpd_nn_sumsq_blocks(t.numel())
                self._gslots.append((off, nb))
                off += nb
            self._gparts = DeviceArray(off)
        for t, (off, nb) in zip(self.gw + self.gb, self._gslots):
            _ck(L.mpd_nn_sumsq(_p(t), C.c_void_p(self._gparts.ptr + 4 * off), t.numel(), s))
        return float(np.sqrt(self._gparts.download().astype(np.float64).sum()))

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8, clip=None):
        """One Adam step; with clip, gradients whose global norm exceeds clip
        are scaled down to it (returns the norm, or None without clip)."""
        L, s = nn_lib(), _stream()
        self.step += 1
        norm = None
        gscale = 1.0
        if clip is not None:
            norm = self.grad_norm()
            if norm > clip:
                gscale = clip / norm
        if getattr(self, "_adam_tab", None) is None:
            # Every weight and bias tensor in one table (one launch per step).
            ts = [(self.w[k], self.wb[k], self.gw[k], self.mw[k], self.vw[k]) for k in range(len(self.layers))]
            ts += [(self.b[k], None, self.gb[k], self.mb[k], self.vb[k]) for k in range(len(self.layers))]
            n = len(ts)
            cols = [(_vp * n)(*[(t[j].ptr if t[j] is not None else None) for t in ts]) for j in range(5)]
            sizes = (C.c_longlong * n)(*[t[0].numel() for t in ts])
            self._adam_tab = (n, cols, sizes)
        n, cols, sizes = self._adam_tab
        _ck(L.mpd_nn_adam_multi(n, *cols, sizes, lr, b1, b2, eps, self.step, gscale, s))
        return norm

    def mse(self, y, t, dy):
        """Mean squared error (host float) and dL/dy into dy (may be None)."""
        L, s = nn_lib(), _stream()
        n = y.numel()
        _ck(L.mpd_nn_mse(_p(y), _p(t), _p(dy) if dy is not None else None, _p(self.loss_part), n, s))
        nb = L.mpd_nn_mse_blocks(n)
        return float(self.loss_part.download()[:nb].astype(np.float64).sum()) / n


# ---------------------------------------------------------------- features

def scale_params(params):
    """Parameters to [-1, 1] over the sweep ranges."""
    return 2.0 * (np.asarray(params, np.float64) - PARAM_LO) / (PARAM_HI - PARAM_LO) - 1.0


def features(params, grid_r, L_max, M_max):
    """(N, L_max, M_max+1, 8) float32 input features of N cases: the scaled
    parameters, 2 z - 1, 2 frac - 1 and (r - 0.5) / 0.3 at every interior
    cell; grid_r: (N, L_max+2, M_max+3) radii of each case's grid."""
    N = len(params)
    Mn = M_max + 1
    sp = scale_params(params)
    out = np.empty((N, L_max, Mn, N_FEATURES), np.float32)
    z = (np.arange(L_max) + 0.0) / L_max
    frac = np.arange(Mn) / M_max
    out[..., :5] = sp[:, None, None, :]
    out[..., 5] = (2 * z - 1)[None, :, None]
    out[..., 6] = (2 * frac - 1)[None, None, :]
    out[..., 7] = (np.asarray(grid_r)[:, 1:L_max + 1, 1:Mn + 1] - 0.5) / 0.3
    return out


def baseline_fields(params, grid_r, L_max, M_max, gamma):
    """(N, L_max, Mn, 8) physical fields of the uniform inflow state each
    case starts from cold (rho_in, u_z_in, B_phi = amp / r, p = 0.5): the
    network predicts the departure of the steady state from it."""
    params = np.asarray(params, np.float64)
    N = len(params)
    Mn = M_max + 1
    r = np.asarray(grid_r)[:, 1:L_max + 1, 1:Mn + 1]
    out = np.zeros((N, L_max, Mn, 8))
    out[..., 0] = params[:, 0, None, None]
    out[..., 1] = params[:, 1, None, None]
    out[..., 4] = 0.5 / ((gamma - 1.0) * params[:, 0, None, None])
    out[..., 7] = params[:, 2, None, None] / r
    return out


def fields_from_phys(phys, L_max, M_max):
    """(L_max, Mn, 8) interior physical fields (OUT_FIELDS order) from the
    solver's 10 physical arrays."""
    Mn = M_max + 1
    return np.stack([phys[PHYS_INDEX[f]][1:L_max + 1, 1:Mn + 1] for f in OUT_FIELDS], axis=-1)


def u0_from_fields(f, r, rows, cols):
    """The conserved arrays u0_1..u0_8 (rows x cols, ghost ring zero) of
    interior physical fields f (L_max, Mn, 8): RebuildUFromPhysical
    (include/physics_utils.hpp:15-24)."""
    L_max, Mn = f.shape[0], f.shape[1]
    rho, vz, vr, vphi, e, hz, hr, hphi = (f[..., k] for k in range(8))
    rr = r[1:L_max + 1, 1:Mn + 1]
    vals = [rho * rr, rho * vz * rr, rho * vr * rr, rho * vphi * rr, rho * e * rr, hphi, hz * rr, hr * rr]
    out = []
    for v in vals:
        a = np.zeros((rows, cols))
        a[1:L_max + 1, 1:Mn + 1] = v
        out.append(a)
    return out


def time_training(H, W, B, steps):
    """Mean milliseconds of one training step (forward, MSE, backward, Adam)
    of the network at batch B on an H x W grid, random data, after up to 10
    untimed steps."""
    import time
    net = UNet(H, W, seed=0)
    rng = np.random.default_rng(0)
    x = DeviceArray.from_numpy(rng.normal(size=(B, net.H, net.W, N_FEATURES)), "bf16")
    t = DeviceArray.from_numpy(rng.normal(size=(B, net.H, net.W, len(OUT_FIELDS))))
    dy = DeviceArray((B, net.H, net.W, len(OUT_FIELDS)))
    warm = min(10, steps - 1)
    for it in range(steps):
        if it == warm:
            _ck(nn_lib().mpd_nn_synchronize())
            t0 = time.perf_counter()
        y = net.forward(x, B)
        net.mse(y, t, dy)
        net.backward(dy, B)
        net.adam_step(1e-3)
    _ck(nn_lib().mpd_nn_synchronize())
    return (time.perf_counter() - t0) / (steps - warm) * 1e3
