"""The two relaxation drivers of the oracle against the every-step residual history they
share (64x64x4 axisymmetric spinning punctures, zero guess):

- main.c's loop (main.c:136-185): residual of y_n before each step; the resident solver stops at
  the first log10 L2 <= target;
- the Einstein Toolkit thorn's loop (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:
  181-229): step first, check the state after n+1 steps when n % k == 0 or n == max_iterations,
  stop on log10 L2 < target (strict), end after max_iterations + 1 steps otherwise.

The expected stop of the thorn's loop is read off the every-step history of a separate run, with
cases chosen so that each plausible slip (checking before the step, counting n instead of n+1,
<= for <, the wrong cadence, an early or late end at max_iterations) gives another step."""
import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import config

CFG = config.axisymmetric_spinning()


@pytest.fixture(scope="module")
def history():
    O = oracle.OracleProblem(CFG)
    st = O.new_state()
    n, hist = O.relax_every_step(st, -7.0, 3000)
    assert hist[-1][1] <= -7.0 and n == hist[-1][0]
    return np.array([h[1] for h in hist])  # l[m] = log10 L2 after m steps


def thorn_expected(l, target, k, max_iterations):
    """First m >= 1 with (m - 1) % k == 0 or m == max_iterations + 1 and l[m] < target."""
    for m in range(1, max_iterations + 2):
        if ((m - 1) % k == 0 or m == max_iterations + 1) and l[m] < target:
            return m, True
    return max_iterations + 1, False


def test_every_step_rule_stops_at_first_value_at_or_below_target(history):
    l = history
    m0 = int(np.argmax(l <= -6.5))
    O = oracle.OracleProblem(CFG)
    n, hist = O.relax_every_step(O.new_state(), float(l[m0]), 3000)  # target hit exactly: <= stops
    assert n == m0


def test_thorn_rule_strict_and_cadence(history):
    l = history
    # a target equal to a new lowest value: the every-step rule (<=) stops there, the thorn's
    # strict < does not
    m0 = int(np.argmax(l <= -6.0))
    assert m0 > 1 and np.all(l[:m0] > l[m0])
    T = float(l[m0])
    for k in (1, 7, 10):
        O = oracle.OracleProblem(CFG)
        st = O.new_state()
        steps, conv, hist = O.relax_et_thorn(st, T, 3000, k)
        want, wconv = thorn_expected(l, T, k, 3000)
        assert conv and wconv and steps == want, (k, steps, want)
        assert steps != m0
        # checks after steps 1, 1+k, 1+2k, ... only
        assert all((m - 1) % k == 0 for m, _ in hist)
        assert [v for _, v in hist] == [float(l[m]) for m, _ in hist]
    # the cadence matters for this target
    assert thorn_expected(l, T, 1, 3000)[0] != thorn_expected(l, T, 7, 3000)[0]


def test_thorn_rule_ends_after_max_iterations_plus_one(history):
    O = oracle.OracleProblem(CFG)
    st = O.new_state()
    steps, conv, hist = O.relax_et_thorn(st, -30.0, 20, 7)
    assert (steps, conv) == (21, False)
    assert [m for m, _ in hist] == [1, 8, 15, 21]  # n = 0, 7, 14 and n == max_iterations
    ref = O.new_state()
    for _ in range(21):
        O.mol_step(ref)
    assert np.array_equal(st["y_n"], ref["y_n"])
    # info_output_freq = 0 means every step
    O2 = oracle.OracleProblem(CFG)
    s2, c2, h2 = O2.relax_et_thorn(O2.new_state(), -30.0, 3, 0)
    assert s2 == 4 and [m for m, _ in h2] == [1, 2, 3, 4]
