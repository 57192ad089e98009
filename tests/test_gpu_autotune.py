"""The tile autotuner: times every admissible (tile_i0, planes) shape over whole RK4 steps on
the configuration's own grid, leaves the configuration's state as it found it, caches the
choice per grid shape in JSON, and the solver runs with the shape it chose."""
import json

import numpy as np
import pytest

from nrpyelliptic_amd import config, runtime
from nrpyelliptic_amd.autotune import CANDIDATES, TileTuner

pytestmark = pytest.mark.gpu


def test_tuner_times_candidates_and_keeps_state(tmp_path):
    P = runtime.Problem(config.unequal_mass_q4(Nxx=(96, 64, 8)))
    y = np.random.default_rng(3).normal(size=2 * P.ntot) * 1e-3
    P.set("y_n_gfs", y)
    cache = tmp_path / "tiles.json"
    tuner = TileTuner(cache_path=str(cache), steps=20)
    us = tuner.measure(P)
    print("us per launch:", {f"{k[0]}x{k[1]}": round(v, 2) for k, v in us.items()}, flush=True)
    assert set(us) == set(CANDIDATES) and all(0 < v < 1e4 for v in us.values())
    assert np.array_equal(P.get("y_n_gfs"), y)
    tx, pl = tuner.choose(P)
    assert us[(tx, pl)] <= min(us.values()) * 1.25  # re-timed inside choose(); within noise of the best
    table = json.loads(cache.read_text())
    assert table["96x64x8_fp64"]["tile_i0"] == tx and table["96x64x8_fp64"]["planes"] == pl
    # a second tuner reads the cache instead of timing again
    assert TileTuner(cache_path=str(cache)).choose(P) == (tx, pl)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S = P.solver(tile_i0=tx, planes=pl, log10_target=-6.0, max_steps=3000)
    assert S.solve().converged
    S.close()
    P.close()
