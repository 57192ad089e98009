"""Tile-width autotuner for the fused RK4 substep kernel.

The fused kernel puts i0 across the lanes of a 256-thread block (tile TX x 256/TX) and gives
each thread one or two i2 planes; which shape keeps HBM busiest depends on N0 (ragged last
tile, BC-tile share) and on the number of blocks per launch. This module times the fused
substep (stage 1 with the residual check, the production launch) for each candidate shape on
the configuration's own grid and caches the winner per (N0, N1, N2, precision) in a JSON
file, so a batch of configurations is grouped by shape without re-timing shapes it has
already seen.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Sequence

from .runtime import Problem, Solver

CANDIDATES = ((16, 1), (32, 1), (64, 1), (32, 2), (64, 2))  # (tile_i0, planes)
_DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_cache.json")


def _key(prob: Problem, precision: int) -> str:
    p = prob.gd.params
    return f"{p.Nxx0}x{p.Nxx1}x{p.Nxx2}_fp{precision}"


class TileTuner:
    """Chooses the tile width of the fused substep kernel per grid shape.

    `cache_path=None` keeps the table in memory only."""

    def __init__(self, cache_path: Optional[str] = _DEFAULT_CACHE, candidates: Sequence[int] = CANDIDATES,
                 launches: int = 200):
        self.cache_path = cache_path
        self.candidates = tuple(candidates)
        self.launches = launches
        self.table: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def measure(self, prob: Problem, precision: int = 64) -> Dict[tuple, float]:
        """Microseconds per fused substep launch for every candidate shape."""
        out = {}
        for tx, pl in self.candidates:
            s = Solver(prob, precision=precision, tile_i0=tx, planes=pl)
            try:
                out[(tx, pl)] = s.time_substep(0, self.launches)
            finally:
                s.close()
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
