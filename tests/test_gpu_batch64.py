"""The batched workload at its full size: the 64 puncture configurations of seed 42 (q
log-uniform in [1,8], D in [6,20], Newtonian momenta x U(0.9,1.1), spins |chi| < 0.8; grids
mixing Nxx0 in {96,128,160}, Nxx1 in {96,128}, Nxx2 in {8,16}) relaxed together to
log10 L2 <= -10 with tile shapes from the autotuner. Configurations whose relaxation diverges at
the reference's CFL factor 0.7 (the reference's C diverges on them too; tests/test_oracle_cfl.py)
are relaxed again from the zero guess with the wave speed at CFL 0.5 (runtime.relax_batch).
Every configuration converges; the members that fall back are the ones the reference's C diverges
on, and each stops at the reference's own step at CFL 0.5; members are spot-checked against their
own single-configuration solves (bit-identical states)."""
import os

import numpy as np
import pytest

from nrpyelliptic_amd import config, runtime
from nrpyelliptic_amd.autotune import TileTuner

pytestmark = pytest.mark.gpu


def test_all_64_seed42_configurations_converge():
    cfgs = config.random_batch(64, seed=42)
    probs = [runtime.Problem(c) for c in cfgs]
    try:
        tiles = TileTuner(cache_path=None).choose_all(probs)
        res, sec, info = runtime.relax_batch(probs, tiles=tiles, log10_target=-10.0, max_steps=30000)
        fell = [k for k in range(64) if info[k]["fallback"]]
        print(f"64 configurations in {sec:.2f} s; fallback at CFL {runtime.FALLBACK_CFL}: {fell}", flush=True)
        assert all(r.converged for r in res), [(k, runtime.STATUS[r.status]) for k, r in enumerate(res)]
        assert all(r.log10_l2 <= -10.0 for r in res)
        assert all(probs[k].cfl == runtime.FALLBACK_CFL for k in fell)
        assert 0 < len(fell) < 16
        # the reference's C diverges on exactly these members and, at the fallback CFL, stops at
        # the same step as the batched solver (tests/golden/cfl_fallback_seed42.npz)
        fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfl_fallback_seed42.npz"))
        assert fell == [int(k) for k in fx["indices"]]
        for k, n in zip(fx["indices"], fx["steps_fallback"]):
            assert res[int(k)].steps == int(n), (int(k), res[int(k)].steps, int(n))
        # spot checks: two members that fell back, two that did not
        plain = [k for k in range(64) if not info[k]["fallback"]]
        for k in fell[:2] + plain[:2]:
            got = probs[k].get("y_n_gfs")
            Q = runtime.Problem(cfgs[k])
            try:
                if info[k]["fallback"]:
                    Q.set_wave_speed_cfl(runtime.FALLBACK_CFL)
                S = Q.solver(log10_target=-10.0, max_steps=30000, check_every=1,
                             tile_i0=tiles[k][0], planes=tiles[k][1])
                r = S.solve()
                S.close()
                assert r.converged and r.steps == res[k].steps, (k, r.steps, res[k].steps)
                assert np.array_equal(Q.get("y_n_gfs"), got), k
            finally:
                Q.close()
    finally:
        for P in probs:
            P.close()
