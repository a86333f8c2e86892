"""The batched path with its launch options: the count-kernel autotuner and
the graph-captured tail (IDB_FLAG_AUTOTUNE / IDB_FLAG_GRAPH) must leave every
result unchanged, across repeated runs (graph replays and in-place updates),
different inputs in one context, and the device-side union / fuse choice."""
from __future__ import annotations

import numpy as np
import pytest

from immunodbscan_amd import _lib, configs, engine, synth
from oracle import oracle as O

BOTH = _lib.IDB_FLAG_AUTOTUNE | _lib.IDB_FLAG_GRAPH
FIELDS = ("clone_id", "core", "copies", "uniques", "counts")


def oracle_of(rep, cl, fast=False):
    return O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming), fast=fast)


def same(a, b, tag):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), (tag, k)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [_lib.IDB_FLAG_AUTOTUNE, _lib.IDB_FLAG_GRAPH, BOTH])
def test_flags_leave_results_unchanged(gpu_ctx, flags):
    ctx = _lib.Context(0)
    cases = [
        ("c1", synth.generate(configs.c1_params()), configs.C1),
        ("c2-20k", synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000)), configs.C2),
        ("c5-10k", synth.generate(configs.c5_params(n_rows=10_000, n_seeds=100)), configs.C5),
    ]
    for tag, rep, cl in cases:
        print(f"case {tag} flags {flags}", flush=True)
        plain = engine.dbscan(rep, cl, ctx=ctx)
        for it in range(3):  # capture, then update / replay in place
            got = engine.dbscan(rep, cl, ctx=ctx, flags=flags)
            same(got, plain, f"{tag} run {it}")
            assert got.stats["graph_used"] == (1 if flags & _lib.IDB_FLAG_GRAPH else 0)
            if flags & _lib.IDB_FLAG_AUTOTUNE:
                assert got.stats["tuned_wpb"] in (1, 2, 4) and got.stats["tuned_grid_mul"] in (1, 2, 4, 8)
    ctx.close()


@pytest.mark.gpu
def test_graph_fuse_fallback_on_device(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    exp = oracle_of(rep, configs.C2, fast=True)
    got = engine.dbscan(rep, configs.C2, edge_capacity=16, flags=BOTH)
    assert got.stats["fuse_from_edges"] == 0 and got.stats["graph_used"] == 1
    for k in FIELDS:
        assert np.array_equal(getattr(got, k), exp[k]), k
    ok = engine.dbscan(rep, configs.C2, flags=BOTH)
    assert ok.stats["fuse_from_edges"] == 1
    same(ok, got, "edges vs fuse")


@pytest.mark.gpu
def test_c3_batch_graph_autotune_equals_separate(gpu_ctx):
    reps = configs.c3_repertoires(n_reps=64, scale=0.05)
    cat = synth.concat(reps)
    got = engine.dbscan_batch(cat, configs.C3, flags=BOTH)
    again = engine.dbscan_batch(cat, configs.C3, flags=BOTH)
    same(got, again, "replay")
    assert got.stats["graph_used"] == 1
    for r, rep in enumerate(reps):
        s, e = int(cat.rep_start[r]), int(cat.rep_start[r + 1])
        exp = oracle_of(rep, configs.C3)
        for k in FIELDS:
            assert np.array_equal(getattr(got, k)[s:e], exp[k]), (r, k)
        assert got.n_uniques[r] == exp["n_uniques"]


@pytest.mark.gpu
def test_small_pair_capacity_same_results_and_statistics(gpu_ctx):
    """A neighbour-pair capacity far below the pairs found sends cluster growth
    to the fuse pass over the tiles: same results and the same count-pass
    statistics as with the pair list."""
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    exp = oracle_of(rep, configs.C2, fast=True)
    big = engine.dbscan(rep, configs.C2)
    small = engine.dbscan(rep, configs.C2, edge_capacity=2048)
    assert big.stats["neighbour_pairs"] > 2048 and big.stats["fuse_from_edges"] == 1
    assert small.stats["fuse_from_edges"] == 0
    for k in FIELDS:
        assert np.array_equal(getattr(big, k), exp[k]), k
        assert np.array_equal(getattr(small, k), exp[k]), k
    for k in ("gated_pairs", "neighbour_pairs", "myers_pairs", "myers_letters", "gated_letters"):
        assert big.stats[k] == small.stats[k], k
