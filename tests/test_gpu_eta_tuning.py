"""The damping-strength tuner (autotune.EtaTuner) on the 128x128x16 grid. Calibrated on a
configuration that is not the one measured, it allows the network's start a larger eta than
the configured one (that start carries less low-frequency error); on the bench problem the
tuned solve converges in at most half the steps of the reference path (zero guess, configured
eta) and its u is at least as close to a solve converged to 1e-14 as the reference path's.
The calibration leaves the problem's eta and state as they were, and a second request is
served from the table. The batched solver at the tuned damping converges on the same
configurations as from the zero guess, in under half the steps, to the same u."""
import os

import numpy as np
import pytest

from nrpyelliptic_amd import config, learned, runtime
from nrpyelliptic_amd.autotune import EtaTuner
from nrpyelliptic_amd.mlp import InitialGuessMLP

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "nrpyelliptic_amd", "data", "mlp_initial_guess.npz")


@pytest.fixture(scope="module")
def tuned():
    net = InitialGuessMLP(seed=7)
    net.load(WEIGHTS)
    calib = runtime.Problem(config.random_batch(1, seed=901, grids=False, Nxx=(128, 128, 16))[0])
    eta0 = calib.gd.params.eta_damping
    state0 = np.random.default_rng(1).uniform(-1e-3, 1e-3, 2 * calib.ntot)
    calib.set("y_n_gfs", state0)
    tuner = EtaTuner(cache_path=None)
    seed = lambda Q: learned.seed_with_fallback(net, Q)
    zero = lambda Q: Q.set("y_n_gfs", np.zeros(2 * Q.ntot))
    f_nn = tuner.choose(calib, seed, "nn")
    f_zero = tuner.choose(calib, zero, "zero")
    print(f"eta tuner: factors nn {f_nn} zero {f_zero}; table {tuner.table}", flush=True)
    assert calib.gd.params.eta_damping == eta0
    assert np.array_equal(calib.get("y_n_gfs"), state0)
    n = tuner.measured
    assert tuner.choose(calib, seed, "nn") == f_nn and tuner.measured == n  # from the table
    calib.close()
    yield net, f_nn, f_zero


def test_tuned_factors_follow_the_damping_picture(tuned):
    _, f_nn, f_zero = tuned
    assert f_nn > 1.0 and f_nn >= f_zero >= 1.0


def test_tuned_solve_is_faster_and_no_less_accurate(tuned):
    net, f_nn, _ = tuned
    P = runtime.Problem(config.equal_mass_boosted())
    m = np.zeros(P.Ng[::-1], dtype=bool)
    m[5:-5, 5:-5, 5:-5] = True
    m = m.ravel()
    eta0 = P.gd.params.eta_damping

    def solve(eta, nn, target):
        if nn:
            used, eta_conf = learned.seed_with_tuned_damping(net, P, eta / eta0)
            assert used and eta_conf == eta0 and abs(P.gd.params.eta_damping - eta) <= 1e-12 * eta
        else:
            P.gd.params.eta_damping = eta
            P.set("y_n_gfs", np.zeros(2 * P.ntot))
        S = P.solver(log10_target=target, max_steps=40000)
        r = S.solve()
        S.close()
        P.gd.params.eta_damping = eta0
        assert r.converged
        return r.steps, P.get("y_n_gfs")[: P.ntot][m].copy()

    _, u_star = solve(eta0, False, -14.0)
    n_ref, u_ref = solve(eta0, False, -10.0)
    n_nn, u_nn = solve(f_nn * eta0, True, -10.0)
    e_ref, e_nn = np.max(np.abs(u_ref - u_star)), np.max(np.abs(u_nn - u_star))
    print(f"bench problem: reference path {n_ref} steps, error {e_ref:.2e}; NN start at eta {f_nn * eta0:g} "
          f"{n_nn} steps, error {e_nn:.2e}", flush=True)
    assert n_ref == 3144 and n_nn <= 0.5 * n_ref
    assert e_nn <= e_ref
    P.close()


def test_batched_solver_with_tuned_damping(tuned):
    """Four held-out configurations relaxed together: from the network's guess at the tuned
    damping (each configuration its own eta) the same configurations converge as from the
    zero guess, each in under half the steps, and each member's u is at least as close to its
    own solve converged to 1e-14 as the zero guess's (the reference path's) u."""
    net, f_nn, f_zero = tuned
    cfgs = config.random_batch(4, seed=900, grids=False, Nxx=(128, 128, 16))
    probs = [runtime.Problem(c) for c in cfgs]
    eta0 = [P.gd.params.eta_damping for P in probs]
    m = np.zeros(probs[0].Ng[::-1], dtype=bool)
    m[5:-5, 5:-5, 5:-5] = True
    m = m.ravel()
    B = runtime.BatchSolver(probs, log10_target=-10.0, max_steps=20000)
    rz, _ = B.solve()
    B.close()
    u_zero = [P.get("y_n_gfs")[: P.ntot].copy() for P in probs]
    u_star = []
    for P in probs:
        P.zero_guess()
        S = P.solver(log10_target=-14.0, max_steps=40000)
        try:
            u_star.append(P.get("y_n_gfs")[: P.ntot].copy() if S.solve().converged else None)
        except runtime.NrpyellError:
            u_star.append(None)
        S.close()
    for P in probs:
        learned.seed_with_tuned_damping(net, P, f_nn, f_zero)
    B = runtime.BatchSolver(probs, log10_target=-10.0, max_steps=20000)
    rn, _ = B.solve()
    B.close()
    print("batched steps zero / tuned:", [(a.steps, b.steps) for a, b in zip(rz, rn)], flush=True)
    for P, e, a, b, u, us in zip(probs, eta0, rz, rn, u_zero, u_star):
        P.gd.params.eta_damping = e
        assert bool(a.converged) == bool(b.converged)
        if a.converged:
            assert b.steps <= 0.5 * a.steps
            u_nn = P.get("y_n_gfs")[: P.ntot]
            assert np.max(np.abs(u_nn - u)) <= 1e-7
            assert us is not None
            e_ref, e_nn = np.max(np.abs((u - us)[m])), np.max(np.abs((u_nn - us)[m]))
            print(f"  {P.cfg.name}: error zero guess {e_ref:.2e}, tuned network start {e_nn:.2e}", flush=True)
            assert e_nn <= e_ref
        P.close()
