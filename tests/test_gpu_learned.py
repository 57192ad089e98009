"""The trained initial-guess network (nrpyelliptic_amd/data/mlp_initial_guess.npz: 6 x 256
SiLU coordinate MLP on the BF16 MFMA kernels, trained by nrpyelliptic_amd.learned on the FP64
solutions of the 4096 configurations drawn with seed 1000) against the plain method on held-out
configurations (seed 900), and the fallback rule (zero guess whenever the network's guess has the
larger initial residual)."""
import os

import numpy as np
import pytest

from nrpyelliptic_amd import config, learned, runtime
from nrpyelliptic_amd.mlp import InitialGuessMLP

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "nrpyelliptic_amd", "data", "mlp_initial_guess.npz")


def test_network_guess_beats_zero_guess_on_held_out_configurations():
    net = InitialGuessMLP(seed=7)
    net.load(WEIGHTS)
    gains = []
    for cfg in config.random_batch(3, seed=900, grids=False, Nxx=(128, 128, 16)):
        P = runtime.Problem(cfg)
        S = P.solver(log10_target=-10.0, max_steps=10000)
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        rz = S.solve()
        u_zero = P.get("y_n_gfs")[: P.ntot]
        used, r0, r1 = learned.seed_with_fallback(net, P)
        rn = S.solve()
        assert used and r1 < r0 - 0.1, (cfg.name, r0, r1)  # lower initial residual
        assert rz.converged and rn.converged and rn.log10_l2 <= -10.0
        assert rn.steps < rz.steps, (cfg.name, rz.steps, rn.steps)  # fewer steps to the same threshold
        # both stop at L2 <= 1e-10 on different trajectories: their remaining errors differ
        assert np.max(np.abs(P.get("y_n_gfs")[: P.ntot] - u_zero)) <= 1e-7
        gains.append(rz.steps / rn.steps)
        S.close()
        P.close()
    print("steps ratio zero/NN:", [round(g, 3) for g in gains])
    assert np.mean(gains) > 1.05


def test_fallback_keeps_the_zero_guess_when_the_network_is_worse():
    net = InitialGuessMLP(seed=3)
    p = net.get_params()
    p[-1] = 8.0  # output bias: t ~ 8, u = e^t / (50 (r+b)) is thousands of times the solution
    net.set_params(p)
    P = runtime.Problem(config.random_batch(1, seed=900, grids=False, Nxx=(64, 64, 8))[0])
    used, r0, r1 = learned.seed_with_fallback(net, P)
    assert not used and r1 >= r0
    assert not np.any(P.get("y_n_gfs"))  # the zero guess is back in place
    P.close()


def test_tuned_damping_falls_back_with_the_zero_guess():
    """A rejected network guess leaves u = v = 0 and the zero guess's damping."""
    net = InitialGuessMLP(seed=3)
    p = net.get_params()
    p[-1] = 8.0
    net.set_params(p)
    P = runtime.Problem(config.random_batch(1, seed=900, grids=False, Nxx=(64, 64, 8))[0])
    eta0 = P.gd.params.eta_damping
    used, eta_conf = learned.seed_with_tuned_damping(net, P, 2.8, 1.0)
    assert not used and eta_conf == eta0 and P.gd.params.eta_damping == eta0
    assert not np.any(P.get("y_n_gfs"))
    P.close()


def test_evaluate_with_tuned_damping():
    """learned.evaluate with the damping tuned to each start (the factors EtaTuner gives on the
    128x128x16 grid): the NN-seeded mixed solve of a held-out configuration takes under half
    the zero guess's steps and lands on the same u, and the configured eta is left in place."""
    net = InitialGuessMLP(seed=7)
    net.load(WEIGHTS)
    cfg = config.random_batch(1, seed=900, grids=False, Nxx=(128, 128, 16))[0]
    rows = learned.evaluate(net, [cfg], precision=32, eta_factors=(2.8, 1.0), log=lambda s: None)
    r = rows[0]
    print("tuned evaluate:", r, flush=True)
    assert r["converged"] and r["nn_used"] and r["nn_steps"] < 0.5 * r["zero_steps"]
    assert r["max_du"] <= 1e-7


def test_network_start_on_32_held_out_configurations_batched():
    """32 held-out configurations (seed 900, 128x128x16) relaxed together from the zero guess and
    from the network's guess (configured eta, learned.evaluate_batched): the same configurations
    converge, the network's guess is kept on nearly all of them, it needs fewer steps in total, and
    the two converged states agree to the level the residual target bounds them."""
    net = InitialGuessMLP(seed=7)
    net.load(WEIGHTS)
    out = learned.evaluate_batched(net, config.random_batch(32, seed=900, grids=False, Nxx=(128, 128, 16)),
                                   1.0, 1.0, log=lambda s: None)
    print("batched held-out:", out, flush=True)
    assert out["converged_zero"] == out["converged_nn"] == 32
    assert out["nn_guess_used"] >= 28
    assert out["speedup_steps"] > 1.05
    assert out["max_du_nn_vs_zero"] <= 1e-7
