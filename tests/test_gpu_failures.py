"""Failure paths of the solvers, each forced from the host: the step cap, NaN in the state
(single solve, mixed precision, one member of a batch) and a configuration whose relaxation
diverges (the reference's C diverges on it as well). Every case ends with an error code or a
per-configuration report, never a hang, and leaves the other configurations' results as a
solve without the bad member gives them."""
import numpy as np
import pytest

from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu


def _poison(P):
    """NaN at one interior point of u (and its value in v left finite)."""
    y = P.get("y_n_gfs")
    G0, G1, G2 = P.Ng
    i0, i1, i2 = 5 + (G0 - 10) // 3, 5 + (G1 - 10) // 2, 5 + (G2 - 10) // 2
    y[i0 + G0 * (i1 + G1 * i2)] = np.nan
    P.set("y_n_gfs", y)


def test_step_cap_reports_not_converged():
    P = runtime.Problem(config.axisymmetric_spinning())
    S = P.solver(log10_target=-10.0, max_steps=50)
    r = S.solve()
    print(f"capped solve: steps {r.steps} log10 {r.log10_l2:.3f} converged {r.converged}", flush=True)
    assert not r.converged and 50 <= r.steps <= 100 and np.isfinite(r.log10_l2) and r.log10_l2 > -8
    S.close()
    P.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_nan_in_state_stops_with_enan(precision):
    P = runtime.Problem(config.axisymmetric_spinning())
    _poison(P)
    S = P.solver(log10_target=-8.0, max_steps=2000, precision=precision)
    print(f"NaN-poisoned solve, precision {precision}", flush=True)
    with pytest.raises(runtime.NrpyellError, match="ENAN"):
        S.solve()
    S.close()
    P.close()


def test_nan_member_of_batch_is_reported_alone():
    cfgs = [config.axisymmetric_spinning(), config.axisymmetric_spinning(), config.unequal_mass_q4(Nxx=(64, 64, 8))]
    probs = [runtime.Problem(c) for c in cfgs]
    # reference result of the healthy members, solved one at a time
    singles = []
    for k in (0, 2):
        S = probs[k].solver(log10_target=-8.0, max_steps=3000)
        r = S.solve()
        singles.append((r.steps, probs[k].get("y_n_gfs")))
        S.close()
        probs[k].set("y_n_gfs", np.zeros(2 * probs[k].ntot))
    _poison(probs[1])
    B = runtime.BatchSolver(probs, log10_target=-8.0, max_steps=3000)
    res, _ = B.solve()
    print("batch with a NaN member:", [(bool(r.converged), r.steps) for r in res], flush=True)
    assert res[0].converged and res[2].converged and not res[1].converged
    for k, (steps, y) in zip((0, 2), singles):
        assert res[k].steps == steps
        assert np.array_equal(probs[k].get("y_n_gfs"), y)
    B.close()
    for P in probs:
        P.close()


def test_diverging_configuration_stops_early():
    """Seed 42's second configuration (128x128x16 grid family) diverges in the reference's C too:
    the solver ends it once its residual is 4 decades above its lowest value, well before the cap."""
    cfg = config.random_batch(4, seed=42)[1]
    P = runtime.Problem(cfg)
    S = P.solver(log10_target=-10.0, max_steps=20000)
    r = S.solve()
    steps, hist = S.history()
    print(f"diverging solve: steps {r.steps} last log10 {r.log10_l2:.2f} lowest {np.min(hist):.2f}", flush=True)
    assert not r.converged and r.steps < 20000
    assert r.log10_l2 > np.min(hist) + 4.0 - 1e-9
    S.close()
    P.close()
