"""Compact digests of a solver state, for parity checks at sizes whose full
state is too large to keep as a fixture (test infrastructure only).

A digest holds, for each of the 8 conserved and 10 physical arrays
((L_max+2) x (M_max+3) float64), a 64-bit BLAKE2b hash of every block of
`block` consecutive rows (bit-exact equality, localised to a block), plus a
strided subsample of the conserved arrays (relative error when bits differ).
"""
import hashlib

import numpy as np

BLOCK_ROWS = 16
SUB_ROWS = 32   # subsample: about this many rows ...
SUB_COLS = 32   # ... and columns per conserved array


def block_hashes(a, block=BLOCK_ROWS):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((a.shape[0] + block - 1) // block, dtype=np.uint64)
    for i, r0 in enumerate(range(0, a.shape[0], block)):
        h = hashlib.blake2b(a[r0:r0 + block].tobytes(), digest_size=8).digest()
        out[i] = np.frombuffer(h, dtype=np.uint64)[0]
    return out


def sub_index(rows, cols):
    rs = max(1, rows // SUB_ROWS)
    cs = max(1, cols // SUB_COLS)
    return slice(0, rows, rs), slice(0, cols, cs)


def make_digest(u0, phys):
    """u0: 8 arrays, phys: 10 arrays."""
    rows, cols = u0[0].shape
    ri, ci = sub_index(rows, cols)
    return dict(u0_hash=np.stack([block_hashes(a) for a in u0]),
                phys_hash=np.stack([block_hashes(a) for a in phys]),
                u0_sub=np.stack([np.asarray(a)[ri, ci] for a in u0]))


def compare(d, ref):
    """Returns a dict: bit-exact flags, the differing blocks and the
    subsample's per-field relative L-inf error."""
    du = d["u0_hash"] != ref["u0_hash"]
    dp = d["phys_hash"] != ref["phys_hash"]
    rel = []
    for k in range(ref["u0_sub"].shape[0]):
        a, b = d["u0_sub"][k], ref["u0_sub"][k]
        scale = np.max(np.abs(b))
        rel.append(float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b))))
    return dict(u0_exact=not du.any(), phys_exact=not dp.any(),
                u0_bad_blocks=[(int(k), int(b)) for k, b in np.argwhere(du)][:20],
                phys_bad_blocks=[(int(k), int(b)) for k, b in np.argwhere(dp)][:20],
                u0_sub_rel=rel)
