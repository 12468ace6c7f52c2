"""GPU parity at the workload sizes themselves (marked slow, all collected).

* The memory-bound channel at 1600 x 640 for its full 10,000 steps and at
  6400 x 2560 for 1000 steps: the final state against digests of the
  oracle's final state (tests/golden/make_digests.py; BLAKE2b hashes of
  every 16-row block of the 8 conserved and 10 physical arrays, plus a
  strided subsample for a relative error), t, dt and step count bit-exact.
* The uneven batched sweep: 64 instances with nz in {200, 400, 800, 1600},
  nr = 0.4 nz, seeded inflow parameters, convergence checks every 100 steps,
  300 steps in one batch.  Every instance equals its own single-solver run
  bit for bit (state, t, dt, SolutionChange); one instance of each size
  equals the oracle bit for bit.
"""
import os

import numpy as np
import pytest

import _oracle as O
import cases
import digest
from parity import describe_mismatch, run_oracle, same_bits

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _digest_case(mpd, name, L, M, steps):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run tests/golden/make_digests.py {name}")
    z = np.load(path)
    text = cases.channel_config(L, M, geometry="condiv", noise=1e-3, max_steps=steps)
    print(f"{name}: {L}x{M}, {steps} steps", flush=True)
    with mpd.Solver(text) as sol:
        res = sol.run()
        u, p = sol.download()
        integ = sol.integrals()
        flux, thrust = sol.outlet()
    d = digest.make_digest(u, p)
    del u, p
    cmp = digest.compare(d, z)
    print({k: cmp[k] for k in ("u0_exact", "phys_exact", "u0_bad_blocks")}, flush=True)
    print("subsample relative L-inf per conserved field:", cmp["u0_sub_rel"], flush=True)
    assert res["steps"] == steps == int(z["steps"])
    assert res["t"] == float(z["t"]) and res["dt"] == float(z["dt"]), (res, float(z["t"]), float(z["dt"]))
    assert max(cmp["u0_sub_rel"]) <= 1e-12
    assert cmp["u0_exact"], f"conserved arrays differ in blocks (field, 16-row block) {cmp['u0_bad_blocks']}"
    assert cmp["phys_exact"], f"physical arrays differ in blocks {cmp['phys_bad_blocks']}"
    assert flux == float(z["mass_flux"]) and thrust == float(z["thrust"])
    # Integrals: the reference sums serially by rows, the device in a tree.
    assert np.all(np.abs(integ - z["integrals"]) <= 1e-12 * np.abs(z["integrals"]) + 1e-300)


def test_condiv_1600x640_10000_steps_matches_oracle_digest(mpd):
    _digest_case(mpd, "digest_condiv_1600x640_10000", 1600, 640, 10000)


def test_condiv_6400x2560_1000_steps_matches_oracle_digest(mpd):
    _digest_case(mpd, "digest_condiv_6400x2560_1000", 6400, 2560, 1000)


def _uneven_configs(n=64, seed=12345, steps=300):
    import sweep
    rng = np.random.default_rng(seed)
    sizes, cfgs = [], []
    for _ in range(n):
        nz = int(rng.choice([200, 400, 800, 1600]))
        q = [rng.uniform(0.5, 2.0), rng.uniform(0.05, 0.3), rng.uniform(0.05, 0.5), 0.05, 0.1]
        sizes.append(nz)
        # A threshold no state reaches: the checks run, nothing stops.
        cfgs.append(sweep.case_config(q, nz, int(nz * 0.4), threshold=1e-300, check_frequency=100,
                                      max_steps=steps))
    return sizes, cfgs


def test_uneven_batch_64_instances(mpd):
    steps = 300
    sizes, cfgs = _uneven_configs(steps=steps)
    print("sizes:", {z: sizes.count(z) for z in (200, 400, 800, 1600)}, flush=True)
    with mpd.Batch(cfgs) as b:
        print(b.autotune(3), flush=True)
        res = b.run()
        firsts = {z: sizes.index(z) for z in (200, 400, 800, 1600)}
        for i, cfg in enumerate(cfgs):
            assert res[i]["steps"] == steps and res[i]["total_steps"] == steps, (i, res[i])
            ub, pb = b.instance(i).download()
            assert all(np.all(np.isfinite(a)) for a in ub), i
            with mpd.Solver(cfg) as sol:
                rs = sol.run()
                us, ps = sol.download()
            assert (rs["t"], rs["dt"], rs["steps"]) == (res[i]["t"], res[i]["dt"], res[i]["steps"]), (i, rs, res[i])
            # The same block partition of the SolutionChange sum in both.
            assert rs["change"] == res[i]["change"] and rs["change"] > 0, (i, rs["change"], res[i]["change"])
            for k in range(8):
                assert same_bits(ub[k], us[k]), f"instance {i} u0_{k + 1}: {describe_mismatch(ub[k], us[k])}"
            for k in range(10):
                assert same_bits(pb[k], ps[k]), f"instance {i} {O.PHYS_NAMES[k]}"
            if i in firsts.values():
                print(f"instance {i} (nz {sizes[i]}) against the oracle", flush=True)
                o, ores = run_oracle(cfg, steps)
                assert ores["t"] == res[i]["t"] and ores["dt"] == res[i]["dt"], (ores, res[i])
                for k in range(8):
                    assert same_bits(ub[k], o.u0(k)), f"u0_{k + 1}: {describe_mismatch(ub[k], o.u0(k))}"
                for k in range(10):
                    assert same_bits(pb[k], o.phys(k)), O.PHYS_NAMES[k]
                # SolutionChange sums in another order in the oracle.
                assert abs(ores["change"] - res[i]["change"]) <= 1e-12 * ores["change"]
                del o
