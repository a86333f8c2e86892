"""The reference's run log, rebuilt from a run's statistics.

The reference writes its messages and stage timers through BenchmarkSingleton
(cpp/BenchmarkSingleton.hpp:33-115) as a two-column TSV: header
"Timer/Message<TAB>Value"; a message row is ("At <s> sec", text) with the
seconds since logging began; a timer row is (stage name, "<s> sec").  Numbers
use std::to_string's six decimals.  DBSCAN emits, in this order
(cuda/PSDBSCAN.hcu:986-997, 1035-1038, 1113, 1198, 1279-1292, 1300-1354):

  settings (epsilon, minPts, the Hamming switch), the number of sequences, the
  duplicate / tree-reduction / copies timers, the number of uniques, the
  core-point timer, the number of core points, the grow-clusters timer, the
  union-find depth histogram (when asked for), the cluster-tree timer, the
  numbers of clusters and noise points and the integer noise percentage, the
  renumbering timer and the whole-run timer.

The "Dispatching Kernel ..." lines of the reference describe its own launch
grids and have no counterpart here.  The depth histogram describes this
build's union-find forest, whose shape (like the reference's) depends on the
linking order, not on the clustering.
"""
from __future__ import annotations

import time

from ._lib import STAGE_NAMES

HEADER = "Timer/Message\tValue"


def _sec(x: float) -> str:
    return f"{x:.6f} sec"


class RunLog:
    """Rows in BenchmarkSingleton's format; messages are stamped with the
    seconds since the log was created, timers carry their own duration."""

    def __init__(self) -> None:
        self.t0 = time.perf_counter()
        self.rows: list[tuple[str, str]] = []

    def message(self, text: str) -> None:
        self.rows.append(("At " + _sec(time.perf_counter() - self.t0), text))

    def timer(self, name: str, seconds: float) -> None:
        self.rows.append((name, _sec(seconds)))

    def text(self) -> str:
        return "\n".join([HEADER] + [f"{a}\t{b}" for a, b in self.rows]) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.text())


def log_dbscan(log: RunLog, n_rows: int, eps: float, min_pts: int, use_hamming: bool, stats: dict,
               tree_stats: bool = False) -> RunLog:
    """Appends one clustering run's rows in the reference's order, from the
    stats dict of engine.dbscan (stage_ms holds the eight reference timers)."""
    ms = stats["stage_ms"]
    t = lambda name: log.timer(name, ms[name] * 1e-3)  # noqa: E731
    log.message(f"DBSCAN settings: epsilon = {eps:.6f}")
    log.message(f"DBSCAN settings: minPts = {min_pts}")
    if use_hamming:
        log.message("DBSCAN settings: using Hamming distance measure instead of Levenshtein distance")
    log.message(f"Total number of sequences = {n_rows}")
    for name in STAGE_NAMES[:3]:
        t(name)
    log.message(f"Number of uniques = {stats['n_uniques']}")
    t(STAGE_NAMES[3])
    log.message(f"Number of core points = {stats['n_cores']}")
    t(STAGE_NAMES[4])
    if tree_stats:
        log.message("Union find tree depth structure")
        for i, c in enumerate(stats["tree_depth"]):
            log.message(f"Number of trees with depth {i}: {c}")
    t(STAGE_NAMES[5])
    log.message(f"Total number of clusters = {stats['n_clusters']}")
    log.message(f"Number of noise points = {stats['n_noise']}")
    # unsigned 32-bit arithmetic, as in the reference
    log.message(f"Percentage of noise points = {((100 * stats['n_noise']) & 0xFFFFFFFF) // max(n_rows, 1)}%")
    t(STAGE_NAMES[6])
    t(STAGE_NAMES[7])
    return log
