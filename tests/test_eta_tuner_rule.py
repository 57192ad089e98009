"""The damping tuner's selection rule (autotune.EtaTuner.choose), on synthetic calibration
results (no GPU): fewest steps among the factors whose error is at most margin x the reference
path's error; ties go to the smaller factor; 1.0 when no factor qualifies or the calibration
failed; a faster but less accurate factor is never chosen; results are cached per (grid, start,
target) and written to the JSON table."""
import json
import types

from nrpyelliptic_amd import autotune


def _prob(n=(128, 128, 16)):
    p = types.SimpleNamespace(Nxx0=n[0], Nxx1=n[1], Nxx2=n[2], AMAX=1e6, SINHWAA=0.07)
    return types.SimpleNamespace(gd=types.SimpleNamespace(params=p))


def _tuner(result, **kw):
    t = autotune.EtaTuner(cache_path=None, **kw)
    calls = []

    def measure(prob, start):
        calls.append(prob)
        return result
    t.measure = measure
    return t, calls


def test_fastest_factor_within_the_accuracy_margin():
    res = {"reference_steps": 3144, "reference_error": 1e-8,
           "factors": {1.0: (2857, 2e-9), 2.0: (1467, 3e-9), 2.8: (1032, 4.7e-9), 4.0: (774, 2.5e-7)}}
    t, _ = _tuner(res)
    assert t.choose(_prob(), None, "nn") == 2.8  # 4.0 is faster but 25x less accurate


def test_margin_boundary_and_ties_go_to_the_smaller_factor():
    res = {"reference_steps": 100, "reference_error": 1e-8,
           "factors": {1.0: (90, 5e-9), 1.4: (80, 5.0000001e-9), 2.0: (90, 1e-9)}}
    t, _ = _tuner(res)
    # 1.0 sits exactly on the 0.5 margin (accepted), 1.4 just above it; 1.0 and 2.0 tie at 90
    assert t.choose(_prob(), None, "x") == 1.0
    t2, _ = _tuner(res, margin=0.6)
    assert t2.choose(_prob(), None, "x") == 1.4


def test_no_qualifying_factor_or_failed_calibration_keeps_the_configured_eta():
    t, _ = _tuner({"reference_steps": 100, "reference_error": 1e-8, "factors": {2.0: (50, 1e-7)}})
    assert t.choose(_prob(), None, "zero") == 1.0
    t2, _ = _tuner({})
    assert t2.choose(_prob(), None, "zero") == 1.0


def test_cached_per_grid_start_and_target(tmp_path):
    res = {"reference_steps": 100, "reference_error": 1e-8, "factors": {2.0: (50, 1e-9)}}
    path = str(tmp_path / "eta.json")
    t, calls = _tuner(res)
    t.cache_path = path
    P = _prob()
    assert t.choose(P, None, "nn") == 2.0 and t.choose(P, None, "nn") == 2.0 and len(calls) == 1
    t.choose(P, None, "zero")
    t.choose(_prob((96, 96, 8)), None, "nn")
    assert len(calls) == 3
    table = json.load(open(path))
    assert len(table) == 3 and all("factor" in v and "reference_error" in v for v in table.values())
    t3 = autotune.EtaTuner(cache_path=path)  # a new tuner reads the table back
    assert t3.choose(P, None, "nn") == 2.0 and t3.measured == 0


# ---- ScheduleTuner: fewest steps until u is within tol of u* ----

def _stuner(result, **kw):
    t = autotune.ScheduleTuner(cache_path=None, **kw)
    calls = []

    def measure(prob, start):
        calls.append(prob)
        return result
    t.measure = measure
    return t, calls


def test_schedule_with_fewest_steps_to_the_accuracy_bar():
    res = {(1.0, -9.0, 1.0): (-12.5, 4887), (2.8, -9.0, 2.8): (-14.0, 2855),
           (2.8, -9.0, 1.4): (-11.5, 2005), (2.8, -9.0, 1.0): (-12.0, 2776)}
    t, _ = _stuner(res)
    got = t.choose(_prob(), None, "nn")
    assert got == {"schedule": (2.8, -9.0, 1.4), "target": -11.5, "steps": 2005}


def test_schedule_ties_go_to_the_earlier_candidate_and_failure_keeps_the_configured_eta():
    a, b = autotune.SCHEDULES[1], autotune.SCHEDULES[5]
    t, _ = _stuner({b: (-12.0, 100), a: (-12.5, 100)})
    assert t.choose(_prob(), None, "x")["schedule"] == a
    t2, _ = _stuner({})
    assert t2.choose(_prob(), None, "x") == {"schedule": (1.0, -9.0, 1.0), "target": None, "steps": None}


def test_schedule_cached_per_grid_start_tolerance_and_precision(tmp_path):
    res = {(2.0, -9.0, 2.0): (-12.5, 2292)}
    path = str(tmp_path / "sched.json")
    t, calls = _stuner(res)
    t.cache_path = path
    P = _prob()
    assert t.choose(P, None, "nn")["steps"] == 2292 and t.choose(P, None, "nn")["steps"] == 2292
    assert len(calls) == 1
    t.choose(P, None, "zero")
    t.tol = 1e-9
    t.choose(P, None, "nn")
    assert len(calls) == 3
    table = json.load(open(path))
    assert len(table) == 3 and all(v["schedule"] == [2.0, -9.0, 2.0] for v in table.values())
    t2 = autotune.ScheduleTuner(cache_path=path)
    t2.measure = lambda prob, start: (_ for _ in ()).throw(AssertionError("re-measured"))
    assert t2.choose(P, None, "nn")["schedule"] == (2.0, -9.0, 2.0)


def test_deepest_fp32_switch_with_no_more_steps_within_the_margin():
    res = {"reference_steps": 3144, "reference_error": 2e-8,
           "factors": {1.0: (2900, 3e-9), 2.8: (1059, 8e-9)},
           "switches": {-6.0: (1059, 8e-9), -7.0: (1059, 9.5e-9), -8.0: (1059, 1.5e-8), -9.0: (1328, 2e-9)}}
    t, _ = _tuner(res, precision=32, fp32_switches=(-6.0, -7.0, -8.0, -9.0))
    P = _prob()
    assert t.choose(P, None, "nn") == 2.8
    # -8 is over the 0.5 x 2e-8 margin, -9 takes more steps: -7
    assert t.fp32_switch_for(P, "nn") == -7.0
    t2, _ = _tuner(dict(res, switches={-6.0: (1059, 8e-9), -7.0: (1100, 5e-9)}), precision=32,
                   fp32_switches=(-6.0, -7.0))
    assert t2.choose(P, None, "nn") == 2.8 and t2.fp32_switch_for(P, "nn") == -6.0
    t3, _ = _tuner({}, precision=32, fp32_switches=(-6.0, -7.0))
    assert t3.choose(P, None, "nn") == 1.0 and t3.fp32_switch_for(P, "nn") == -6.0
