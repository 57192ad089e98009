"""Autotuners of the relaxation: the fused kernel's tile shape (TileTuner), the damping
strength eta (EtaTuner) and, for an absolute accuracy bar, the damping schedule (ScheduleTuner).

Tile width of the fused RK4 substep kernel:

The fused kernel puts i0 across the lanes of a 256-thread block (tile TX x 256/TX) and gives
each thread one or two i2 planes; which shape keeps HBM busiest depends on N0 (ragged last
tile, BC-tile share) and on the number of blocks per launch. This module times whole
graph-replayed RK4 steps (the four fused launches with the residual check, as the solver runs
them) for each candidate shape on the configuration's own grid, restoring the configuration's
state afterwards, and caches the winner per (N0, N1, N2, precision) in a JSON file, so a batch
of configurations is grouped by shape without re-timing shapes it has already seen. (Timing
one stage launched back to back misranks variants whose effect shows only when the next stage
reads what this one wrote.)
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict, Iterable, Optional, Sequence

import numpy as np

from .runtime import NrpyellError, Problem, Solver

CANDIDATES = ((16, 1), (32, 1), (64, 1), (32, 2), (64, 2))  # (tile_i0, planes)
_DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_cache.json")


def _key(prob: Problem, precision: int) -> str:
    p = prob.gd.params
    return f"{p.Nxx0}x{p.Nxx1}x{p.Nxx2}_fp{precision}"


class TileTuner:
    """Chooses the tile width of the fused substep kernel per grid shape.

    `cache_path=None` keeps the table in memory only."""

    def __init__(self, cache_path: Optional[str] = _DEFAULT_CACHE, candidates: Sequence[int] = CANDIDATES,
                 steps: int = 50):
        self.cache_path = cache_path
        self.candidates = tuple(candidates)
        self.steps = steps
        self.table: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def measure(self, prob: Problem, precision: int = 64) -> Dict[tuple, float]:
        """Microseconds per fused substep launch, over whole RK4 steps, for every candidate
        shape the grid admits. The steps are timed in FP64 (the FP32 phase runs the same
        kernel on half the bytes); the configuration's state is restored afterwards."""
        out = {}
        saved = prob.get("y_n_gfs")
        try:
            for tx, pl in self.candidates:
                try:
                    s = Solver(prob, tile_i0=tx, planes=pl, steps_per_graph=self.steps)
                except NrpyellError:
                    continue  # shape not available on this grid
                try:
                    s.timed_steps(self.steps)  # warm-up (graph instantiation, caches)
                    ms, _ = s.timed_steps(self.steps)
                    out[(tx, pl)] = 1e3 * ms / (4 * self.steps)
                finally:
                    s.close()
        finally:
            prob.set("y_n_gfs", saved)
        return out

    def choose(self, prob: Problem, precision: int = 64) -> tuple:
        """(tile_i0, planes) of the fastest shape for this grid."""
        k = _key(prob, precision)
        if k not in self.table:
            us = self.measure(prob, precision)
            best = min(us, key=us.get)
            self.table[k] = {"tile_i0": best[0], "planes": best[1],
                             "us_per_launch": {f"{t[0]}x{t[1]}": v for t, v in us.items()}}
            self._save()
        return int(self.table[k]["tile_i0"]), int(self.table[k]["planes"])

    def choose_all(self, problems: Iterable[Problem], precision: int = 64):
        return [self.choose(p, precision) for p in problems]

    def _save(self):
        if not self.cache_path:
            return
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)


# ---------------------------------------------------------------------------------------
# Damping strength of the relaxation
# ---------------------------------------------------------------------------------------
ETA_FACTORS = (1.0, 1.4, 2.0, 2.8, 4.0, 5.6)  # multiples of the configured eta


def _eta_key(prob: Problem, start: str, target: float, precision: int) -> str:
    p = prob.gd.params
    return f"{p.Nxx0}x{p.Nxx1}x{p.Nxx2}_A{p.AMAX:g}_w{p.SINHWAA:g}_{start}_t{target:g}_p{precision}"


class EtaTuner:
    """Chooses the damping strength eta of the hyperbolic relaxation per grid and start, at no
    loss of accuracy.

    In u_tt + eta u_t = c^2 (lap u + ...) a spatial mode of eigenvalue lambda decays at rate
    eta/2 while eta < 2 c sqrt(lambda) (underdamped) and at c^2 lambda / eta above it. A larger
    eta therefore reaches the residual target in fewer steps, but past some point only because
    the residual (masked to r < integration radius) no longer sees the slow, long-wavelength
    error that is left: the converged u gets worse. A start that carries less low-frequency
    error, such as the network's guess, stays accurate up to a larger eta.

    `choose(prob, start, name)` calibrates on `prob` (a configuration of the grid, not the one
    to be measured): u* = a solve from the zero guess at the configured eta to target - 4; the
    reference error e_ref = max |u - u*| of the zero-guess solve at the configured eta to the
    target (the reference driver's path); then, for every factor of the configured eta, a
    solve from the initial state `start(prob)` writes to the target, in `precision` (32 for the
    FP32 -> FP64 path the network's start runs on). The factor with the fewest steps among those
    whose error is at most `margin` x e_ref is cached per (grid, start, target, precision), 1.0 if
    none qualifies; configurations of the grid with other punctures
    and foci reuse it. The problem's eta and state are restored afterwards. `cache_path=None`
    keeps the table in memory only.

    With precision 32 and more than one level in `fp32_switches`, the chosen factor's solve is
    repeated with the FP32 -> FP64 switch at each level (log10 L2; the first level is the one the
    factors were measured at): the deepest level whose solve takes no more steps and stays within
    the same accuracy margin is cached too (fp32_switch_for), since an FP32 step costs less than an
    FP64 one and the FP32 state's rounding is what a deeper switch risks."""

    def __init__(self, cache_path: Optional[str] = None, factors: Sequence[float] = ETA_FACTORS,
                 target: float = -10.0, max_steps: int = 40000, margin: float = 0.5, precision: int = 64,
                 fp32_switches: Sequence[float] = (-6.0,)):
        self.fp32_switches = tuple(float(x) for x in fp32_switches)
        self.precision = precision  # of the solves from `start` (the path the factor will serve)
        self.cache_path = cache_path
        self.factors = tuple(factors)
        self.target = target
        self.max_steps = max_steps
        self.margin = margin
        self.measured = 0  # calibrations run (not served from the table)
        self.table: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def _solve(self, prob: Problem, eta: float, start: Callable[[Problem], object], target: float,
               precision: int = 64, fp32_switch: Optional[float] = None):
        """(steps, interior u) of one solve, or None if it does not converge."""
        prob.gd.params.eta_damping = eta
        start(prob)  # after eta: a seed writes v = eta u
        kw = {} if fp32_switch is None else {"fp32_switch_log10": fp32_switch}
        s = Solver(prob, log10_target=target, max_steps=self.max_steps, precision=precision, **kw)
        try:
            r = s.solve()
        except NrpyellError:
            return None  # NaN at this eta
        finally:
            s.close()
        if not r.converged:
            return None
        return int(r.steps), prob.get("y_n_gfs")[: prob.ntot][_interior(prob)]

    def measure(self, prob: Problem, start: Callable[[Problem], object]) -> dict:
        """Reference error and (steps, error) per converging factor."""
        eta0 = prob.gd.params.eta_damping
        saved = prob.get("y_n_gfs")
        zero = lambda q: q.set("y_n_gfs", np.zeros(2 * q.ntot))
        try:
            star = self._solve(prob, eta0, zero, self.target - 4.0)
            ref = self._solve(prob, eta0, zero, self.target)
            if star is None or ref is None:
                return {}
            u_star = star[1]
            out = {"reference_steps": ref[0], "reference_error": float(np.max(np.abs(ref[1] - u_star))),
                   "factors": {}}
            sw0 = self.fp32_switches[0] if self.precision == 32 else None
            for f in self.factors:
                got = self._solve(prob, f * eta0, start, self.target, self.precision, sw0)
                if got is not None:
                    out["factors"][f] = (got[0], float(np.max(np.abs(got[1] - u_star))))
            if self.precision == 32 and len(self.fp32_switches) > 1:
                ok = self._qualifying(out)
                if ok:
                    best = min(ok, key=lambda f: (ok[f], f))
                    out["switches"] = {sw0: out["factors"][best]}
                    for sw in self.fp32_switches[1:]:
                        got = self._solve(prob, best * eta0, start, self.target, self.precision, sw)
                        if got is not None:
                            out["switches"][sw] = (got[0], float(np.max(np.abs(got[1] - u_star))))
        finally:
            prob.gd.params.eta_damping = eta0
            prob.set("y_n_gfs", saved)
        self.measured += 1
        return out

    def _qualifying(self, m: dict) -> dict:
        return {f: v[0] for f, v in m["factors"].items() if v[1] <= self.margin * m["reference_error"]}

    def choose(self, prob: Problem, start: Callable[[Problem], object], start_name: str) -> float:
        """Factor of the configured eta: fewest steps at no loss of accuracy (1.0 by default)."""
        k = _eta_key(prob, start_name, self.target, self.precision)
        if k not in self.table:
            m = self.measure(prob, start)
            best = 1.0
            switch = self.fp32_switches[0]
            if m:
                ok = self._qualifying(m)
                if ok:
                    best = min(ok, key=lambda f: (ok[f], f))
                sws = m.get("switches") or {}
                if switch in sws:
                    steps0 = sws[switch][0]
                    good = [sw for sw, v in sws.items() if v[0] <= steps0 and v[1] <= self.margin * m["reference_error"]]
                    switch = min(good)
            self.table[k] = {"factor": best, "fp32_switch": switch,
                             "reference_steps": m.get("reference_steps"), "reference_error": m.get("reference_error"),
                             "factors": {f"{f:g}": {"steps": v[0], "error": v[1]}
                                         for f, v in (m.get("factors") or {}).items()},
                             "fp32_switches": {f"{sw:g}": {"steps": v[0], "error": v[1]}
                                               for sw, v in (m.get("switches") or {}).items()}}
            if self.cache_path:
                tmp = self.cache_path + ".tmp"
                with open(tmp, "w") as fh:
                    json.dump(self.table, fh, indent=1, sort_keys=True)
                os.replace(tmp, self.cache_path)
        return float(self.table[k]["factor"])

    def fp32_switch_for(self, prob: Problem, start_name: str) -> float:
        """The FP32 -> FP64 switch level cached with the factor (after choose; the first level of
        fp32_switches by default)."""
        e = self.table.get(_eta_key(prob, start_name, self.target, self.precision), {})
        return float(e.get("fp32_switch", self.fp32_switches[0]))


# (f_hi, switch log10 L2, f_lo): multiples of the configured eta before and after the switch;
# f_hi == f_lo is one damping throughout
SCHEDULES = ((1.0, -9.0, 1.0), (1.4, -9.0, 1.4), (2.0, -9.0, 2.0), (2.8, -9.0, 2.8),
             (2.8, -9.0, 1.0), (2.8, -9.0, 1.4), (2.8, -9.0, 2.0), (2.0, -9.0, 1.4), (4.0, -9.0, 1.4),
             (4.0, -8.0, 2.0))
ACCURACY_TARGETS = tuple(-10.0 - 0.5 * k for k in range(10))  # -10 .. -14.5


def _schedule_key(prob: Problem, start: str, tol: float, precision: int) -> str:
    p = prob.gd.params
    return f"{p.Nxx0}x{p.Nxx1}x{p.Nxx2}_A{p.AMAX:g}_w{p.SINHWAA:g}_{start}_tol{tol:g}_p{precision}"


class ScheduleTuner:
    """Chooses the damping schedule of a start for accuracy: the fewest RK4 steps until
    max |u - u*| <= tol, with u* the converged solution.

    EtaTuner's constant factor minimises the steps to the residual target at no loss of accuracy
    against the reference path at that target. When the bar is instead an absolute accuracy (u
    within tol of the converged u), the residual target has to go deeper, and there a weaker
    damping late in the relaxation pays: the strong damping that removes the guess's grid-scale
    error fast relaxes the long-wavelength modes (rate c^2 lambda / eta once overdamped) slowly.

    `choose(prob, start, name)` calibrates on `prob` (a configuration of the grid, not the one to
    be measured): u* = a solve from the zero guess at the configured eta to log10 L2 = -14; then
    for each schedule (f_hi, switch, f_lo) of `schedules` the relaxation from `start(prob)` runs at
    f_hi x the configured eta (in `precision`, 32 for the FP32 -> FP64 path) down to the switch,
    continues at f_lo x it (Problem.set_eta_damping keeps u and u_t) and is continued through the
    residual targets -10, -10.5, ... (the state (u, v) carries the trajectory, so each
    continuation is the single longer solve) until the first whose u is within tol of u*. The
    schedule with the fewest steps wins (ties: the earlier in `schedules`); the result --
    schedule, residual target and steps -- is cached per (grid, start, tol, precision), and the
    problem's eta and state are restored. `cache_path=None` keeps the table in memory only."""

    def __init__(self, cache_path: Optional[str] = None, schedules=SCHEDULES, tol: float = 1e-10,
                 targets: Sequence[float] = ACCURACY_TARGETS, max_steps: int = 40000, precision: int = 32):
        self.cache_path = cache_path
        self.schedules = tuple(tuple(float(x) for x in s) for s in schedules)
        self.tol = tol
        self.targets = tuple(targets)
        self.max_steps = max_steps
        self.precision = precision
        self.measured = 0
        self.table: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def _relax(self, prob: Problem, target: float, precision: int):
        s = Solver(prob, log10_target=target, max_steps=self.max_steps, precision=precision)
        try:
            r = s.solve()
        except NrpyellError:
            return None  # NaN
        finally:
            s.close()
        return r if r.converged else None

    def trajectory(self, prob: Problem, start: Callable[[Problem], object], schedule, u_star: np.ndarray,
                   precision: Optional[int] = None):
        """(first residual target whose u is within tol of u*, steps to it), or None; the first
        solve in `precision` (the tuner's by default), the continuations in FP64."""
        f_hi, switch, f_lo = schedule
        eta0 = prob.gd.params.eta_damping
        m = _interior(prob)
        try:
            prob.gd.params.eta_damping = f_hi * eta0
            start(prob)  # after eta: a seed writes v = eta u
            steps, precision = 0, self.precision if precision is None else precision
            if f_lo != f_hi:
                r = self._relax(prob, switch, precision)
                if r is None:
                    return None
                steps, precision = steps + int(r.steps), 64
                prob.set_eta_damping(f_lo * eta0)
            for t in self.targets:
                r = self._relax(prob, t, precision)
                if r is None:
                    return None
                steps, precision = steps + int(r.steps), 64
                err = float(np.max(np.abs(prob.get("y_n_gfs")[: prob.ntot][m] - u_star)))
                if err <= self.tol:
                    return t, steps
            return None
        finally:
            prob.gd.params.eta_damping = eta0

    def measure(self, prob: Problem, start: Callable[[Problem], object]) -> dict:
        eta0 = prob.gd.params.eta_damping
        saved = prob.get("y_n_gfs")
        try:
            prob.zero_guess()
            star = self._relax(prob, -14.0, 64)
            if star is None:
                return {}
            u_star = prob.get("y_n_gfs")[: prob.ntot][_interior(prob)]
            out = {}
            for sch in self.schedules:
                got = self.trajectory(prob, start, sch, u_star)
                if got is not None:
                    out[sch] = got
        finally:
            prob.gd.params.eta_damping = eta0
            prob.set("y_n_gfs", saved)
        self.measured += 1
        return out

    def choose(self, prob: Problem, start: Callable[[Problem], object], start_name: str) -> dict:
        """{"schedule": (f_hi, switch, f_lo), "target": residual target, "steps": steps} on the
        calibration problem; the configured eta throughout (1, -9, 1) when nothing qualifies."""
        k = _schedule_key(prob, start_name, self.tol, self.precision)
        if k not in self.table:
            m = self.measure(prob, start)
            entry = {"schedule": [1.0, -9.0, 1.0], "target": None, "steps": None}
            if m:
                best = min(m, key=lambda sch: (m[sch][1], self.schedules.index(sch)))
                entry = {"schedule": list(best), "target": m[best][0], "steps": m[best][1]}
            entry["measured"] = {f"{a:g}/{b:g}/{c:g}": {"target": v[0], "steps": v[1]} for (a, b, c), v in m.items()}
            self.table[k] = entry
            if self.cache_path:
                tmp = self.cache_path + ".tmp"
                with open(tmp, "w") as fh:
                    json.dump(self.table, fh, indent=1, sort_keys=True)
                os.replace(tmp, self.cache_path)
        e = self.table[k]
        return {"schedule": tuple(e["schedule"]), "target": e["target"], "steps": e["steps"]}


def _interior(prob: Problem) -> np.ndarray:
    G0, G1, G2 = prob.Ng
    m = np.zeros((G2, G1, G0), dtype=bool)
    m[5:-5, 5:-5, 5:-5] = True
    return m.ravel()
