"""Tile-width autotuner for the fused RK4 substep kernel.

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
from typing import Dict, Iterable, Optional, Sequence

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
