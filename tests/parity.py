"""Comparison helpers for the parity tests (test infrastructure only)."""
import numpy as np

import _oracle as O
import cases


def same_bits(a, b):
    """Bit-identical, except that NaN matches NaN whatever its sign/payload
    (NaN payloads depend on operand order, which compilers may swap)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(np.where(na, 0.0, a).view(np.uint64), np.where(nb, 0.0, b).view(np.uint64))


def rel_linf(a, b):
    """max |a - b| / max |b| over the finite entries of b (NaN positions must match)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN positions differ"
    m = np.isfinite(b)
    scale = np.max(np.abs(b[m])) if m.any() else 0.0
    d = np.max(np.abs(a[m] - b[m])) if m.any() else 0.0
    return d / scale if scale > 0 else d


def run_oracle(text, steps=0):
    """Run the oracle's time loop for `steps` steps (0: the config's
    time.max_steps; a config without a cap must pass steps)."""
    case = cases.oracle_case(text)
    n = steps or case["max_steps"]
    assert n > 0 or case["T"] < 1e6, "uncapped oracle run"
    o = O.Oracle(case)
    res = o.run(n)
    return o, res


def describe_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    idx = np.argwhere(bad)
    if len(idx) == 0:
        return "no mismatch"
    l, m = idx[0]
    return f"{len(idx)} cells differ, first at (l={l}, m={m}): {a[l, m]!r} vs {b[l, m]!r}"
