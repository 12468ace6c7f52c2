"""Host-side tests of the warm-start network's C ABI (CPU only): the library
loads and exports every symbol include/mpd_nn.h declares, and the entry points
reject bad arguments before any device work, with the error text naming the
rule."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def nn():
    import surrogate
    return surrogate.nn_lib()


def test_nn_library_exports_every_declared_symbol(nn):
    import surrogate
    hdr = open(os.path.join(REPO, "include", "mpd_nn.h")).read()
    declared = set(re.findall(r"\b(mpd_nn_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(surrogate.NN_EXPORTS), declared ^ set(surrogate.NN_EXPORTS)
    for name in declared:
        assert hasattr(nn, name), name


def _fails(nn, rc, text):
    assert rc != 0
    assert text in nn.mpd_nn_last_error().decode()


# A non-null pointer the argument checks accept; nothing is ever read
# through it (every call below is rejected first).
P = C.c_void_p(16)


def test_fused_entry_points_reject_bad_arguments(nn):
    s = None
    _fails(nn, nn.mpd_nn_conv3x3_fwd_silu(P, P, P, P, None, 1, 4, 4, 8, 8, s), "conv3x3_fwd_silu")
    _fails(nn, nn.mpd_nn_conv3x3_fwd_silu(P, P, P, P, P, 0, 4, 4, 8, 8, s), "conv3x3_fwd_silu")
    _fails(nn, nn.mpd_nn_silu_bwd_bias(P, None, P, P, P, 10, 8, s), "silu_bwd_bias")
    _fails(nn, nn.mpd_nn_silu_bwd_bias(P, P, P, P, P, 0, 8, s), "silu_bwd_bias")
    _fails(nn, nn.mpd_nn_conv3x3_bwd_data_silu(P, P, P, P, P, P, 1, 4, 4, 12, 8, s), "multiples of 8")
    _fails(nn, nn.mpd_nn_conv3x3_bwd_data_silu(P, P, P, P, P, P, 1, 4, 4, 96, 8, s), "96")
    _fails(nn, nn.mpd_nn_conv3x3_bwd_data_silu(P, P, None, P, P, P, 1, 4, 4, 8, 8, s), "conv3x3_bwd_data_silu")
    # Channel counts of the one-pass skip joins.
    _fails(nn, nn.mpd_nn_upcat_fwd(P, 12, P, 8, P, 1, 2, 2, s), "multiples of 8")
    _fails(nn, nn.mpd_nn_upcat_fwd(P, 8, P, 0, P, 1, 2, 2, s), "multiples of 8")
    _fails(nn, nn.mpd_nn_upcat_bwd(P, P, 6, P, 8, 1, 2, 2, s), "multiples of 4")
    _fails(nn, nn.mpd_nn_upcat_bwd(P, P, 8, None, 8, 1, 2, 2, s), "upcat_bwd")
    _fails(nn, nn.mpd_nn_pool_bwd_add(P, P, P, 1, 3, 4, 8, s), "H, W even")
    _fails(nn, nn.mpd_nn_pool_bwd_add(P, P, P, 1, 4, 4, 6, s), "multiple of 4")
    _fails(nn, nn.mpd_nn_pool_silu_bwd_bias(P, P, P, P, P, P, 1, 4, 3, 8, s), "H, W even")
    _fails(nn, nn.mpd_nn_pool_silu_bwd_bias(P, P, None, P, P, P, 1, 4, 4, 8, s), "pool_silu_bwd_bias")
    _fails(nn, nn.mpd_nn_pool_silu_bwd_bias(P, P, P, P, P, P, 64, 512, 512, 256, s), "2^31")
    # 32-bit index limit of the one-pass joins.
    _fails(nn, nn.mpd_nn_upcat_fwd(P, 512, P, 512, P, 64, 512, 512, s), "upcat_fwd")


def test_adam_multi_rejects_bad_arguments(nn):
    n = (C.c_longlong * 2)(4, -1)
    ptrs = (C.c_void_p * 2)(16, 16)
    _fails(nn, nn.mpd_nn_adam_multi(-1, None, None, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, 1, 1.0, None),
           "adam_multi")
    _fails(nn, nn.mpd_nn_adam_multi(2, ptrs, None, ptrs, ptrs, ptrs, n, 1e-3, 0.9, 0.999, 1e-8, 1, 1.0, None),
           "adam_multi")
    n2 = (C.c_longlong * 2)(4, 4)
    _fails(nn, nn.mpd_nn_adam_multi(2, ptrs, None, ptrs, ptrs, ptrs, n2, 1e-3, 0.9, 0.999, 1e-8, 0, 1.0, None),
           "adam_multi")
    nulls = (C.c_void_p * 2)(16, None)
    _fails(nn, nn.mpd_nn_adam_multi(2, ptrs, None, nulls, ptrs, ptrs, n2, 1e-3, 0.9, 0.999, 1e-8, 1, 1.0, None),
           "adam_multi")
