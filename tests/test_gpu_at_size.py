"""Parity at the sizes the benchmark runs.

The benchmark's batched, 2M-row and 20M-row lines switch on code paths that
smaller inputs never take: the library radix sort of the entry order (above
2^18 entries), 1,024-pair chunks of the neighbour-pair list (lists of 2^24
pairs or more), the device-side union over tens of millions of pairs and the
5-word Myers kernel with the 3-gram bound at full scale.  These tests run them
at full or near-full size:

  * 400,000 rows of the 2M-row generator against the bucketed oracle, every
    output bit for bit (crosses both size thresholds);
  * the full 512-repertoire batch (autotuned, graph-captured) against the
    bucketed oracle per repertoire;
  * the full 20M-row nucleotide workload: whole-output invariants, plus exact
    representatives, copies, neighbour counts, core flags and labels of 2,000
    sampled rows, computed on the CPU over their V/J buckets
    (orc_sample_rows);
  * the trained learned filter on the 2M-row workload, against the exact
    path: neighbour-pair recall and ARI >= 0.999.
"""
from __future__ import annotations

import numpy as np
import pytest

from immunodbscan_amd import _lib, configs, engine, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("clone_id", "core", "copies", "uniques", "counts")
BOTH = _lib.IDB_FLAG_AUTOTUNE | _lib.IDB_FLAG_GRAPH


def oparams(cl):
    return O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming)


def assert_fields(got, exp, tag, s=0, e=None):
    for k in FIELDS:
        g = getattr(got, k)[s:e]
        if not np.array_equal(g, exp[k]):
            bad = np.nonzero(g != exp[k])[0]
            raise AssertionError(f"{tag} {k}: {len(bad)} rows differ, first {bad[:5]}")


@pytest.mark.timeout(600)
def test_c4_generator_400k_crosses_large_input_paths(gpu_ctx):
    rep = synth.generate(configs.c4_params(n_rows=400_000))
    print(f"case c4-400k: n={rep.n}", flush=True)
    got = engine.dbscan(rep, configs.C4)
    # both size-dependent paths are on: more than 2^18 entries (library
    # radix sort) and a default pair capacity of 64 per entry >= 2^24 pairs
    # (1,024-pair chunks)
    assert got.stats["bucket_entries"] > (1 << 18)
    assert 64 * got.stats["bucket_entries"] >= (1 << 24)
    assert got.stats["fuse_from_edges"] == 1
    exp = O.dbscan(O.Data.of(rep), oparams(configs.C4), fast=True)
    assert_fields(got, exp, "c4-400k")
    assert got.stats["neighbour_pairs"] == exp["n_edges"]
    assert got.stats["n_clusters"] == exp["n_clusters"] and got.stats["n_noise"] == exp["n_noise"]


@pytest.mark.timeout(600)
def test_c3_full_batch_equals_oracle_per_repertoire(gpu_ctx):
    reps = configs.c3_repertoires()
    cat = synth.concat(reps)
    print(f"case c3-full: {len(reps)} repertoires, n={cat.n}", flush=True)
    ctx = _lib.Context(0)
    got = engine.dbscan_batch(cat, configs.C3, ctx=ctx, flags=BOTH)
    again = engine.dbscan_batch(cat, configs.C3, ctx=ctx, flags=BOTH)  # graph replay
    ctx.close()
    assert got.stats["graph_used"] == 1
    assert got.stats["bucket_entries"] > (1 << 18)
    for k in FIELDS:
        assert np.array_equal(getattr(got, k), getattr(again, k)), k
    edges = 0
    for r, rep in enumerate(reps):
        s, e = int(cat.rep_start[r]), int(cat.rep_start[r + 1])
        exp = O.dbscan(O.Data.of(rep), oparams(configs.C3), fast=True)
        assert_fields(got, exp, f"c3 repertoire {r}", s, e)
        assert got.n_uniques[r] == exp["n_uniques"], r
        edges += exp["n_edges"]
    assert got.stats["neighbour_pairs"] == edges


def canonical_labels_check(cid, core, n_noise):
    """R12 over the whole output: label 0 is noise, core rows are never
    noise, every cluster has a core row and clusters are numbered by
    ascending smallest core row."""
    assert int((cid == 0).sum()) == n_noise
    assert not np.any((core == 1) & (cid == 0))
    k = int(cid.max())
    rows = np.nonzero(core == 1)[0]
    first = np.full(k + 1, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(first, cid[rows].astype(np.int64), rows)
    assert np.all(first[1:] < np.iinfo(np.int64).max), "a cluster without a core row"
    assert np.all(np.diff(first[1:]) > 0), "clusters not numbered by smallest core row"


@pytest.mark.timeout(900)
def test_c5_full_invariants_and_sampled_rows(gpu_ctx):
    rep = synth.generate(configs.c5_params())
    cl = configs.C5
    print(f"case c5-20M: n={rep.n}", flush=True)
    got = engine.dbscan(rep, cl)
    n = rep.n
    st = got.stats
    print(f"c5-20M: {st['n_uniques']} uniques, {st['n_cores']} cores, {st['n_clusters']} clusters", flush=True)
    # whole-output invariants
    assert int(got.copies.sum(dtype=np.uint64)) == n
    assert int((got.copies > 0).sum()) == got.n_uniques
    assert np.array_equal(got.uniques[:got.n_uniques], np.nonzero(got.copies)[0])
    assert int(got.core.sum()) == st["n_cores"]
    assert np.all(got.core[got.copies == 0] == 0)
    assert np.all((got.counts >= cl.min_pts) == (got.core == 1)) or cl.min_pts == 0
    canonical_labels_check(got.clone_id, got.core, st["n_noise"])
    assert int(got.clone_id.max()) == st["n_clusters"]
    # 2,000 sampled rows against the CPU
    xs = np.sort(np.random.default_rng(55).choice(n, 2000, replace=False))
    smp = O.sample_rows(O.Data.of(rep), oparams(cl), xs)
    for s, x in enumerate(xs):
        r = int(smp["rep"][s])
        if r == x:
            assert got.copies[x] == smp["copies"][s], x
            assert got.counts[x] == smp["count"][s], (x, got.counts[x], smp["count"][s])
            assert got.core[x] == (smp["count"][s] >= cl.min_pts), x
            nb = smp["neighbours"][s]
            cores = nb[got.core[nb] == 1]
            if got.core[x]:
                # R9: a core row shares its label with every core neighbour
                assert np.all(got.clone_id[cores] == got.clone_id[x]), x
            elif len(cores):
                # R10: a border row takes its smallest core neighbour's label
                assert got.clone_id[x] == got.clone_id[cores.min()], x
            else:
                assert got.clone_id[x] == 0, x
        else:
            # R4/R11: a duplicate row has no copies, no core flag, and its
            # representative's label
            assert got.copies[x] == 0 and got.core[x] == 0, x
            assert got.clone_id[x] == got.clone_id[r], x


@pytest.mark.timeout(900)
def test_learned_filter_at_2m_rows_recall_and_ari(gpu_ctx):
    """The trained learned filter on the 2M-row workload, judged by the exact
    path: neighbour-pair recall and ARI of at least 0.999.  Training uses
    5M pairs labelled by the exact kernels on a separate repertoire; the
    threshold is tuned on a third."""
    from immunodbscan_amd import learned

    cl = configs.C4
    vocab = learned.vocab_for(synth.LETTERS[synth.AA])
    train = synth.generate(configs.c2_params(n_rows=1_000_000, n_seeds=100_000, seed=41))
    pairs, labels = learned.training_pairs(train, cl, 5_000_000, seed=3)
    print(f"case learned-2M: {len(labels)} training pairs, {labels.mean():.3f} positive", flush=True)
    assert len(labels) >= 4_900_000
    enc = learned.Encoder(vocab, seed=2)
    st = enc.train(train, pairs, labels, learned.TrainParams(epochs=4, batch_pairs=4096, lr=5e-4, margin=0.9))
    print("train", st, flush=True)
    tune = synth.generate(configs.c2_params(n_rows=400_000, n_seeds=40_000, seed=43))
    tau = learned.tune_threshold(enc.embed(tune), learned.neighbour_pairs(tune, cl), recall=0.9998, slack=0.01)
    test = synth.generate(configs.c4_params())
    ex = engine.dbscan(test, cl)
    le = learned.dbscan_learned(test, cl, enc, tau)
    recall = le.stats["neighbour_pairs"] / ex.stats["neighbour_pairs"]
    a = learned.ari(ex.clone_id, le.clone_id)
    print(f"tau {tau:.4f} recall {recall:.5f} ari {a:.5f} proposals {le.stats['gated_pairs']} "
          f"exact gated {ex.stats['gated_pairs']}", flush=True)
    assert recall >= 0.999 and a >= 0.999
    # verified pairs are exact: every pair the learned run keeps is a neighbour
    assert le.stats["neighbour_pairs"] <= ex.stats["neighbour_pairs"]
    assert np.all(le.copies == ex.copies)
