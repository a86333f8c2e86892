"""The reference's run log (cpp/BenchmarkSingleton.hpp, cuda/PSDBSCAN.hcu
messages): row order and format from fixed statistics, and (GPU) the numbers
of a real run against the oracle's clustering."""
from __future__ import annotations

import re

import numpy as np
import pytest

from immunodbscan_amd import _lib
from immunodbscan_amd.runlog import HEADER, RunLog, log_dbscan

STATS = {"n_uniques": 1800, "n_cores": 700, "n_clusters": 90, "n_noise": 333,
         "tree_depth": list(range(51)), "stage_ms": {n: 1.5 * (k + 1) for k, n in enumerate(_lib.STAGE_NAMES)}}


def test_rows_follow_the_reference_order_and_format():
    log = log_dbscan(RunLog(), 2000, 0.15, 3, True, STATS, tree_stats=True)
    lines = log.text().splitlines()
    assert lines[0] == HEADER == "Timer/Message\tValue"
    firsts = [ln.split("\t")[0] for ln in lines[1:]]
    texts = [ln.split("\t")[1] for ln in lines[1:]]
    msgs = [t for f, t in zip(firsts, texts) if f.startswith("At ")]
    assert msgs[:4] == ["DBSCAN settings: epsilon = 0.150000", "DBSCAN settings: minPts = 3",
                        "DBSCAN settings: using Hamming distance measure instead of Levenshtein distance",
                        "Total number of sequences = 2000"]
    assert all(re.fullmatch(r"At \d+\.\d{6} sec", f) for f in firsts if f.startswith("At "))
    timers = [(f, t) for f, t in zip(firsts, texts) if not f.startswith("At ")]
    assert [f for f, _ in timers] == _lib.STAGE_NAMES
    assert timers[0][1] == "0.001500 sec" and timers[7][1] == "0.012000 sec"
    order = [f if not f.startswith("At ") else t for f, t in zip(firsts, texts)]
    i = order.index
    assert i("Determine copies") < i("Number of uniques = 1800") < i("Identify core points") < \
        i("Number of core points = 700") < i("Grow clusters") < i("Union find tree depth structure") < \
        i("Number of trees with depth 50: 50") < i("Reduction of cluster information tree structure") < \
        i("Total number of clusters = 90") < i("Number of noise points = 333") < \
        i("Percentage of noise points = 16%") < i("Re-number clusters and identify noise") < i("Clustering core process")
    plain = log_dbscan(RunLog(), 2000, 0.15, 3, False, STATS).text()
    assert "Hamming" not in plain and "depth" not in plain


def test_noise_percentage_uses_unsigned_32_bit_arithmetic():
    st = dict(STATS, n_noise=50_000_000)
    txt = log_dbscan(RunLog(), 60_000_000, 0.15, 2, False, st).text()
    assert f"Percentage of noise points = {(5_000_000_000 % 2**32) // 60_000_000}%" in txt


@pytest.mark.gpu
def test_numbers_of_a_run_match_the_oracle(gpu_ctx):
    from immunodbscan_amd import configs, engine, synth
    from oracle import oracle as O

    rep = synth.generate(configs.c1_params())
    cl = configs.C1
    got = engine.dbscan(rep, cl, tree_stats=True)
    exp = O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming))
    txt = log_dbscan(RunLog(), rep.n, cl.eps, cl.min_pts, cl.use_hamming, got.stats, tree_stats=True).text()
    assert f"Number of uniques = {exp['n_uniques']}" in txt
    assert f"Number of core points = {int(np.sum(exp['core']))}" in txt
    assert f"Total number of clusters = {int(exp['clone_id'].max())}" in txt
    noise = int(np.sum(exp["clone_id"] == 0))
    assert f"Number of noise points = {noise}" in txt
    assert f"Percentage of noise points = {100 * noise // rep.n}%" in txt
    depth = [ln for ln in txt.splitlines() if "Number of trees with depth" in ln]
    assert len(depth) == 51
