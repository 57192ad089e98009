"""Checkpoint/resume and the residual log: a solve resumed from a saved state continues bit for
bit (an RK4 step reads nothing but y_n); a checkpoint of another configuration is refused; the
solver's residual history is written in main.c's residual_l2_norm.txt format."""
import re

import numpy as np
import pytest

from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu


def test_resume_from_checkpoint_is_bit_exact(tmp_path):
    cfg = config.axisymmetric_spinning()
    P = runtime.Problem(cfg)
    S = P.solver()
    S.steps(300)
    ck = str(tmp_path / "state.npz")
    P.save_state(ck, steps=300)
    S.steps(300)
    straight = P.get("y_n_gfs")
    S.close()
    P.close()
    Q = runtime.Problem(cfg)
    assert Q.load_state(ck) == 300
    T = Q.solver()
    T.steps(300)
    assert np.array_equal(Q.get("y_n_gfs"), straight)
    T.close()
    other = runtime.Problem(config.axisymmetric_spinning(Nxx=(64, 64, 6)))
    with pytest.raises(runtime.NrpyellError, match="checkpoint"):
        other.load_state(ck)
    other.close()
    Q.close()


def test_residual_log_format(tmp_path):
    P = runtime.Problem(config.axisymmetric_spinning())
    S = P.solver(log10_target=-6.0, max_steps=3000)
    r = S.solve()
    path = tmp_path / "residual_l2_norm.txt"
    n = S.write_residual_log(str(path))
    steps, vals = S.history()
    lines = path.read_text().splitlines()
    assert n == len(lines) == len(steps) and n >= r.steps
    pat = re.compile(r"^ {0,5}\d{1,6} [ \d.e+-]{10} -?\d\.\d{17}e[+-]\d{2}$")
    for line, k, v in zip(lines, steps, vals):
        assert pat.match(line), line
        a, b, c = line.split()
        assert int(a) == k and float(b) == float("%10.4e" % (k * P.dt)) and float(c) == v
    S.close()
    P.close()


def test_mixed_precision_log_spans_both_phases():
    """After an FP32 -> FP64 solve the history holds the FP32 phase's checks and then the FP64
    phase's, numbered by the steps taken since the start: steps strictly increase and the log
    ends at the step the solve stopped on."""
    P = runtime.Problem(config.axisymmetric_spinning())
    S = P.solver(precision=32, log10_target=-8.0, max_steps=5000, check_every=10, fp32_switch_log10=-5.0)
    r = S.solve()
    assert r.converged and 0 < r.fp32_steps < r.steps
    steps, vals = S.history()
    assert np.all(np.diff(steps) > 0)
    assert steps[-1] == r.steps and vals[-1] == r.log10_l2
    assert np.any(steps < r.fp32_steps) and np.any(steps > r.fp32_steps)
    S.close()
    P.close()


def test_changed_damping_between_solves_is_used():
    """eta_damping edited in the params between two solves on one handle takes effect (the
    captured graph is rebuilt): the second solve equals a fresh solver's at the new eta."""
    cfg = config.axisymmetric_spinning()
    P = runtime.Problem(cfg)
    S = P.solver(log10_target=-7.0, max_steps=5000)
    S.solve()
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    P.gd.params.eta_damping = 2.0 * cfg.eta
    r2 = S.solve()
    u2 = P.get("y_n_gfs")
    S.close()
    P.close()
    Q = runtime.Problem(cfg)
    Q.gd.params.eta_damping = 2.0 * cfg.eta
    T = Q.solver(log10_target=-7.0, max_steps=5000)
    r3 = T.solve()
    assert r2.steps == r3.steps
    assert np.array_equal(u2, Q.get("y_n_gfs"))
    T.close()
    Q.close()
