"""The resident solver's two stopping rules against the oracle's drivers (64x64x4 axisymmetric
spinning punctures, zero guess): the every-step rule (the residual of the state before each step,
stop at log10 L2 <= target) and the Einstein Toolkit thorn's rule (step first, decide on the
checks after steps 1, 1+k, 1+2k, ... with a strict <, end after max_iterations + 1 steps;
NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:181-229). Same stop step as the
oracle; u to 1e-10."""
import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu

CFG = config.axisymmetric_spinning()
K = 7


@pytest.fixture(scope="module")
def history():
    O = oracle.OracleProblem(CFG)
    _, hist = O.relax_every_step(O.new_state(), -7.0, 3000)
    return np.array([h[1] for h in hist])


def gpu_solve(**opts):
    P = runtime.Problem(CFG)
    S = P.solver(**opts)
    r = S.solve()
    u = P.get("y_n_gfs")
    S.close()
    P.close()
    return r, u


def test_thorn_rule_stops_where_the_oracle_does(history):
    T = -6.0
    l = history
    due = [m for m in range(1, len(l)) if (m - 1) % K == 0]
    stop = next(m for m in due if l[m] < T)
    # the target is far from every value the decision compares (GPU and oracle residuals agree
    # to ~1e-12 in log10), and the two rules stop at different steps here
    assert min(abs(l[m] - T) for m in due if m <= stop) > 1e-6
    first_le = int(np.argmax(l <= T))
    assert first_le != stop
    O = oracle.OracleProblem(CFG)
    st = O.new_state()
    steps, conv, _ = O.relax_et_thorn(st, T, 3000, K)
    assert conv and steps == stop
    print(f"thorn rule: stop after {stop} steps (every-step rule: {first_le})", flush=True)
    r, u = gpu_solve(stop_rule="et_thorn", check_every=K, log10_target=T, max_steps=3000)
    assert r.converged and r.steps == stop and r.log10_l2 < T
    assert np.max(np.abs(u - st["y_n"])) <= 1e-10
    r2, _ = gpu_solve(log10_target=T, max_steps=3000)
    assert r2.converged and r2.steps == first_le


def test_thorn_rule_ends_after_max_iterations_plus_one():
    O = oracle.OracleProblem(CFG)
    st = O.new_state()
    steps, conv, _ = O.relax_et_thorn(st, -30.0, 20, K)
    assert (steps, conv) == (21, False)
    r, u = gpu_solve(stop_rule="et_thorn", check_every=K, log10_target=-30.0, max_steps=20)
    assert not r.converged and r.steps == 21 and runtime.STATUS[r.status] == "max_steps"
    assert np.max(np.abs(u - st["y_n"])) <= 1e-10


def test_thorn_rule_options_are_checked():
    P = runtime.Problem(CFG)
    try:
        with pytest.raises(runtime.NrpyellError, match="precision 64"):
            P.solver(stop_rule="et_thorn", precision=32)
        with pytest.raises(runtime.NrpyellError, match="EVERY_STEP"):
            runtime.BatchSolver([P], stop_rule="et_thorn")
    finally:
        P.close()
