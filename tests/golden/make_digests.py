"""Writes the state digests of tests/golden/digest_*.npz: the oracle
(oracle/liblf_oracle.so, the C restatement of the reference's step) run for
the full step count of a workload whose final state is too large to keep
(see tests/digest.py for the digest).  The GPU tests run the same config
text through the library and compare digests.

Run from the repository root (OMP_NUM_THREADS sets the oracle's threads):
    python tests/golden/make_digests.py [name ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402
import digest  # noqa: E402
from parity import run_oracle  # noqa: E402

# name -> (config text, steps)
DIGESTS = {
    # The 1600 x 640 memory-bound workload for its full 10,000 steps.
    "digest_condiv_1600x640_10000": (cases.channel_config(1600, 640, geometry="condiv", noise=1e-3,
                                                          max_steps=10000), 10000),
    # The 6400 x 2560 workload (the throughput benchmark's) for 1000 steps.
    "digest_condiv_6400x2560_1000": (cases.channel_config(6400, 2560, geometry="condiv", noise=1e-3,
                                                          max_steps=1000), 1000),
}


def make(name):
    text, steps = DIGESTS[name]
    t0 = time.time()
    o, res = run_oracle(text, steps)
    d = digest.make_digest([o.u0(k) for k in range(8)], [o.phys(k) for k in range(10)])
    np.savez_compressed(os.path.join(HERE, name + ".npz"), steps=np.int64(res["steps"]), t=np.float64(res["t"]),
                        dt=np.float64(res["dt"]), integrals=o.integrals(), mass_flux=np.float64(o.mass_flux()),
                        thrust=np.float64(o.thrust()), **d)
    print(f"wrote {name}: {res} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(DIGESTS)):
        make(n)
