"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer and index outputs must be bit-exact: clone ids, core flags, copies,
uniques and the exact copies-weighted neighbour counts.  Small and medium
inputs are checked against the literal all-pairs oracle; the full 200k-row
workload against the bucketed oracle (itself pinned to the literal one).
"""
from __future__ import annotations

import random

import numpy as np
import pytest

from immunodbscan_amd import _lib, configs, engine, synth
from oracle import oracle as O
from tests.test_oracle import CONFIGS, make_rows, random_rows

pytestmark = pytest.mark.gpu


def rep_from_rows(rows) -> synth.Repertoire:
    data, seqs, vb, jb = make_rows(rows)
    return synth.Repertoire(data.seq[: int(data.len.sum())].copy(), data.off, data.len, vb, jb,
                            np.zeros(len(rows), np.uint32))


def oracle_of(rep, cl, fast=False):
    return O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming), fast=fast)


def assert_same(got, exp, tag=""):
    for k in ("copies", "core", "uniques", "clone_id", "counts"):
        g = getattr(got, k)
        if not np.array_equal(g, exp[k]):
            bad = np.where(g != exp[k])[0]
            raise AssertionError(f"{tag} {k}: {len(bad)} rows differ, first {bad[:5]} got {g[bad[:5]]} "
                                 f"want {exp[k][bad[:5]]}")
    assert got.n_uniques == exp["n_uniques"]
    assert got.stats["n_cores"] == exp["n_cores"] and got.stats["n_clusters"] == exp["n_clusters"]
    assert got.stats["n_noise"] == exp["n_noise"]


def check(rep, cl, fast=False, tag="", **kw):
    print(f"case {tag}: n={rep.n} cl={cl}", flush=True)
    got = engine.dbscan(rep, cl, **kw)
    exp = oracle_of(rep, cl, fast=fast)
    assert_same(got, exp, tag)
    if exp["n_uniques"]:
        assert got.stats["neighbour_pairs"] == exp["n_edges"], tag
    return got, exp


# --------------------------------------------------------------------------
# Workloads
def test_c1_full(gpu_ctx):
    rep = synth.generate(configs.c1_params())
    check(rep, configs.C1, tag="c1")


def test_c2_subset_literal(gpu_ctx):
    rep = synth.generate(configs.c2_params()).subset(np.arange(20_000))
    check(rep, configs.C2, tag="c2-20k")


def test_c2_full_bucketed(gpu_ctx):
    rep = synth.generate(configs.c2_params())
    got, exp = check(rep, configs.C2, fast=True, tag="c2-200k")
    assert got.stats["fuse_from_edges"] == 1


def test_c5_generator_small(gpu_ctx):
    rep = synth.generate(configs.c5_params(n_rows=20_000, n_seeds=200))
    check(rep, configs.C5, fast=True, tag="c5-20k")


@pytest.mark.parametrize("cfg", CONFIGS)
def test_random_configs(gpu_ctx, cfg):
    rng = random.Random(40 + CONFIGS.index(cfg))
    rows = random_rows(rng, 3000, b"ACGT", 6, 40, 6, 3, 0.3, p_empty=0.03, families=150)
    cl = configs.Clustering(eps=cfg.get("eps", 0.15), abs_k=cfg.get("abs_k", -1), min_pts=cfg["min_pts"],
                            use_hamming=cfg["use_hamming"])
    check(rep_from_rows(rows), cl, tag=str(cfg))


@pytest.mark.parametrize("alphabet", [b"ACGTN", bytes(range(1, 33)), bytes(range(256)), b"ACDEFGHIKLMNPQRSTVWY"])
@pytest.mark.parametrize("hamming", [False, True])
def test_alphabets(gpu_ctx, alphabet, hamming):
    rng = random.Random(len(alphabet) * 7 + hamming)
    rows = random_rows(rng, 2000, alphabet, 5, 30, 4, 2, 0.2, families=100, edits=1 if hamming else 2)
    cl = configs.Clustering(eps=0.1, abs_k=-1 if hamming else 2, min_pts=2, use_hamming=hamming)
    check(rep_from_rows(rows), cl, tag=f"K={len(alphabet)} ham={hamming}")


# random_rows may insert up to `edits` letters: keep hi + 3 within the
# supported lengths (256 for Levenshtein, 1024 nt for Hamming)
@pytest.mark.parametrize("lo,hi,hamming", [(100, 250, False), (200, 253, False), (33, 70, False), (500, 1000, True),
                                           (60, 64, False), (1, 3, False)])
def test_lengths(gpu_ctx, lo, hi, hamming):
    rng = random.Random(lo + hi)
    rows = random_rows(rng, 1200, b"ACGT", lo, hi, 3, 2, 0.2, families=60, edits=3)
    cl = configs.Clustering(eps=0.02 if hamming else 0.0, abs_k=-1 if hamming else 3, min_pts=2,
                            use_hamming=hamming)
    check(rep_from_rows(rows), cl, tag=f"len {lo}-{hi}")


def test_amino_acid_hamming_long_rows(gpu_ctx):
    """Alphabets of 5..32 letters pack six 5-bit letters per word; rows of up
    to 384 letters fit the 64 words a row may have."""
    rng = random.Random(77)
    rows = random_rows(rng, 1200, b"ACDEFGHIKLMNPQRSTVWY", 250, 380, 3, 2, 0.2, families=60, edits=3)
    rows += [(b"W" * 384, [0], [0]), (b"W" * 383 + b"Y", [0], [0])]
    check(rep_from_rows(rows), configs.Clustering(eps=0.02, abs_k=-1, min_pts=2, use_hamming=True), tag="aa 250-384")
    with pytest.raises(_lib.IdbError) as e:
        engine.dbscan(rep_from_rows([(b"ACDEF" * 77, [0], [0])]), configs.C1)  # 385 letters, 5-letter alphabet
    assert e.value.code == _lib.IDB_ERR_UNSUPPORTED and "384 letters" in str(e.value)


# --------------------------------------------------------------------------
# Edge cases and fallbacks
def test_empty_and_single(gpu_ctx):
    for rows in ([], [("ACGT", [0], [0])], [("", [0], [0])], [("", [0], [0]), ("", [0], [0]), ("A", [0], [0])]):
        rep = rep_from_rows(rows)
        for cl in (configs.C1, configs.C2, configs.Clustering(eps=0.5, min_pts=1)):
            check(rep, cl, tag=f"tiny {len(rows)}")


def test_all_duplicates(gpu_ctx):
    rows = [("CASSLGQGYEQYF", [3], [1])] * 1500 + [("CASSLGQGYEQYW", [3], [1])] * 10
    got, _ = check(rep_from_rows(rows), configs.Clustering(eps=0.0, abs_k=1, min_pts=5), tag="dups")
    assert got.copies[0] == 1500 and got.copies[1500] == 10


def test_forced_hash_collisions(gpu_ctx):
    rng = random.Random(8)
    rows = random_rows(rng, 3000, b"ACGT", 4, 9, 3, 2, 0.3, families=80, edits=1)
    rep = rep_from_rows(rows)
    for bits in (1, 3, 8):
        check(rep, configs.Clustering(eps=0.0, abs_k=1, min_pts=3), tag=f"hash{bits}", debug_hash_bits=bits)


def test_edge_list_overflow_falls_back_to_fuse_pass(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    a = engine.dbscan(rep, configs.C2)
    b = engine.dbscan(rep, configs.C2, edge_capacity=16)
    assert a.stats["fuse_from_edges"] == 1 and b.stats["fuse_from_edges"] == 0
    exp = oracle_of(rep, configs.C2, fast=True)
    assert_same(a, exp, "edges")
    assert_same(b, exp, "fuse pass")
    rep1 = synth.generate(configs.c1_params())
    c = engine.dbscan(rep1, configs.C1, edge_capacity=3)
    assert c.stats["fuse_from_edges"] == 0
    assert_same(c, oracle_of(rep1, configs.C1), "fuse pass hamming")


def test_many_genes_single_bucket_mode(gpu_ctx):
    rng = random.Random(9)
    rows = random_rows(rng, 2500, b"ACGT", 6, 12, 64, 4, 0.0, families=120, edits=1)
    rows = [(s, sorted(set(v + [rng.randrange(64) for _ in range(8)])), j + [rng.randrange(4)]) for s, v, j in rows]
    got, _ = check(rep_from_rows(rows), configs.Clustering(eps=0.0, abs_k=1, min_pts=3), tag="many genes")
    assert got.stats["bucket_entries"] == got.n_uniques  # one bucket per repertoire


def test_gene_widths(gpu_ctx):
    rng = random.Random(10)
    rows = random_rows(rng, 2000, b"ACGT", 6, 12, 100, 40, 0.4, families=100, edits=1)
    check(rep_from_rows(rows), configs.Clustering(eps=0.0, abs_k=1, min_pts=2), tag="vw=4 jw=2")
    rows = [(s, [], j) for s, v, j in rows]
    got, _ = check(rep_from_rows(rows), configs.Clustering(eps=0.0, abs_k=1, min_pts=2), tag="no V genes")
    assert got.stats["neighbour_pairs"] == 0


@pytest.mark.parametrize("eps", [float("nan"), float("inf"), 0.0, 1.0, 3.0, -0.3])
@pytest.mark.parametrize("min_pts", [0, 1, 4, 2**32 - 1])
def test_parameter_extremes(gpu_ctx, eps, min_pts):
    rng = random.Random(12)
    rows = random_rows(rng, 600, b"ACGT", 4, 10, 3, 2, 0.3, families=40, edits=2)
    rep = rep_from_rows(rows)
    for hamming in (False, True):
        check(rep, configs.Clustering(eps=eps, abs_k=-1, min_pts=min_pts, use_hamming=hamming),
              tag=f"eps={eps} minPts={min_pts} ham={hamming}")


def test_unsupported_length_is_a_clean_error(gpu_ctx):
    rows = [("A" * 300, [0], [0]), ("A" * 299, [0], [0])]
    with pytest.raises(_lib.IdbError) as e:
        engine.dbscan(rep_from_rows(rows), configs.C2)
    assert e.value.code == _lib.IDB_ERR_UNSUPPORTED
    # the context stays usable
    check(rep_from_rows(rows[:1]), configs.C1, tag="after error")


def test_row_count_past_the_slot_indexes_is_a_clean_error(gpu_ctx):
    """2^31 rows or more are refused before any input is read: the arrays here hold one row."""
    ctx = _lib.Context(0)
    seq = np.zeros(16, np.uint8)
    off = np.zeros(1, np.uint64)
    length = np.ones(1, np.uint32)
    bits = np.zeros(1, np.int32)
    d = _lib.make_dataset(1 << 31, seq, off, length, bits, 1, bits, 1)
    prm = engine._params(configs.C2, False, 0, 0, 0)
    outs = _lib.make_outputs(*[np.zeros(1, np.uint32) for _ in range(4)])
    print("case n_rows=2^31 with one-row arrays", flush=True)
    assert _lib.lib.idb_dbscan(ctx.handle, d, prm, outs, None) == _lib.IDB_ERR_UNSUPPORTED
    assert "2^31" in ctx.last_error()
    rows = [("CASSLG", [0], [0]), ("CASSLA", [0], [0])]
    check(rep_from_rows(rows), configs.C2, ctx=ctx, tag="after error")
    ctx.close()


# --------------------------------------------------------------------------
# Batch, boundary and device-resident paths
def test_batch_equals_separate_runs(gpu_ctx):
    reps = configs.c3_repertoires(n_reps=24, scale=0.05)
    empty = synth.Repertoire(np.zeros(0, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32),
                             np.zeros((0, reps[0].vw), np.int32), np.zeros((0, reps[0].jw), np.int32),
                             np.zeros(0, np.uint32))
    reps.insert(5, empty)  # an empty repertoire in the middle
    cat = synth.concat(reps)
    got = engine.dbscan_batch(cat, configs.C3)
    for r, rep in enumerate(reps):
        s, e = int(cat.rep_start[r]), int(cat.rep_start[r + 1])
        exp = oracle_of(rep, configs.C3)
        for k in ("clone_id", "core", "copies", "uniques", "counts"):
            assert np.array_equal(getattr(got, k)[s:e], exp[k]), (r, k)
        assert got.n_uniques[r] == exp["n_uniques"]


def test_batch_of_one_equals_single(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=10_000, n_seeds=1_000))
    one = synth.concat([rep])
    a = engine.dbscan(rep, configs.C2)
    b = engine.dbscan_batch(one, configs.C2)
    for k in ("clone_id", "core", "copies", "uniques", "counts"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_interlink_python_boundary(gpu_ctx):
    from immunodbscan_amd.interlink import InterlinkClass

    rng = random.Random(13)
    vh = [ln.decode() for ln in open("tests/golden/VHgenes.txt", "rb").read().split(b"\n") if ln]
    jh = [ln.decode() for ln in open("tests/golden/JHgenes.txt", "rb").read().split(b"\n") if ln]
    rows = random_rows(rng, 1500, b"ACDEFGHIKLMNPQRSTVWY", 8, 20, len(vh), len(jh), 0.3, families=100)
    triples = []
    for s, v, j in rows:
        vcall = ", or ".join(f"Homsap {vh[g][:-1]}*01 F" for g in v)
        jcall = ", or ".join(f"Homsap {jh[g][:-1]}*02 F" for g in j)
        triples.append((vcall, jcall, s.decode()))
    il = InterlinkClass(triples, len(triples), vh, len(vh), jh, len(jh), 0.15, 2, False, True)
    # same rows through the oracle: bitsets from substring search
    vb = O.gene_bitsets([t[0].encode() for t in triples], [g.encode() for g in vh])
    jb = O.gene_bitsets([t[1].encode() for t in triples], [g.encode() for g in jh])
    seqs = [t[2].encode() for t in triples]
    lens = np.array([len(s) for s in seqs], np.uint32)
    off = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64)
    d = O.Data(np.frombuffer(b"".join(seqs), np.uint8), off, lens, vb, jb)
    exp = O.dbscan(d, O.params(0.15, -1, 2, False))
    assert np.array_equal(il.get_clusters_iterator(), exp["clone_id"])
    assert np.array_equal(il.get_cores_iterator(), exp["core"])
    assert np.array_equal(il.get_copies_iterator(), exp["copies"])
    assert np.array_equal(il.get_uniques_iterator(), exp["uniques"])
    assert il.get_n_uniques() == exp["n_uniques"]
    assert il.get_v_bitset_width() == 4
    assert np.array_equal(il.v_gene_container_begin(), vb.reshape(-1))
    assert sum(il.stats.tree_depth) == len(triples)
    # the run log carries the reference's messages with this run's numbers
    txt = il.run_log.text()
    assert f"Number of uniques = {exp['n_uniques']}" in txt
    assert f"Total number of clusters = {int(exp['clone_id'].max())}" in txt
    assert "Union find tree depth structure" in txt


def test_device_resident_path(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=30_000, n_seeds=3_000))
    run = engine.DeviceRun(rep, configs.C2)
    run()
    run()
    exp = oracle_of(rep, configs.C2, fast=True)
    assert_same(run.result(), exp, "device")


def test_repeat_runs_are_identical(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=50_000, n_seeds=5_000))
    a = engine.dbscan(rep, configs.C2)
    for _ in range(3):
        b = engine.dbscan(rep, configs.C2)
        for k in ("clone_id", "core", "copies", "uniques", "counts"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        for k in ("gated_pairs", "neighbour_pairs", "tiles", "bucket_entries"):
            assert a.stats[k] == b.stats[k], k


def test_interlink_cpp_drop_in(gpu_ctx, tmp_path):
    """The C++ InterlinkClass (same declarations as the reference's
    cuda/Interlink.hpp) driven like TableHolder::do_the_clustering."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "interlink_demo")
    assert os.path.exists(exe), "interlink_demo not built"
    rng = random.Random(14)
    vh = [ln.decode() for ln in open("tests/golden/VHgenes.txt", "rb").read().split(b"\n") if ln]
    jh = [ln.decode() for ln in open("tests/golden/JHgenes.txt", "rb").read().split(b"\n") if ln]
    rows = random_rows(rng, 2000, b"ACGT", 30, 60, len(vh), len(jh), 0.2, families=120, edits=2)
    tsv = tmp_path / "rows.tsv"
    with open(tsv, "w") as f:
        for s, v, j in rows:
            f.write(", or ".join(f"Homsap {vh[g][:-1]}*01 F" for g in v) + "\t"
                    + ", or ".join(f"Homsap {jh[g][:-1]}*01 F" for g in j) + "\t" + s.decode() + "\n")
    for eps, minpts, ham in ((0.15, 2, 1), (0.1, 3, 0)):
        runlog = tmp_path / f"runlog_{ham}.tsv"
        out = subprocess.run([exe, str(tsv), "tests/golden/VHgenes.txt", "tests/golden/JHgenes.txt", str(eps),
                              str(minpts), str(ham), str(runlog), "1"], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        lines = [ln for ln in out.stdout.split("\n") if ln and not ln.startswith("#")]
        got = np.array([[int(x) for x in ln.split("\t")] for ln in lines])
        vb = O.gene_bitsets([(", or ".join(f"Homsap {vh[g][:-1]}*01 F" for g in v)).encode() for _, v, _ in rows],
                            [g.encode() for g in vh])
        jb = O.gene_bitsets([(", or ".join(f"Homsap {jh[g][:-1]}*01 F" for g in j)).encode() for _, _, j in rows],
                            [g.encode() for g in jh])
        seqs = [s for s, _, _ in rows]
        lens = np.array([len(s) for s in seqs], np.uint32)
        off = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64)
        exp = O.dbscan(O.Data(np.frombuffer(b"".join(seqs), np.uint8), off, lens, vb, jb),
                       O.params(eps, -1, minpts, bool(ham)))
        assert np.array_equal(got[:, 0], exp["clone_id"])
        assert np.array_equal(got[:, 1], exp["copies"])
        assert np.array_equal(got[:, 2], exp["core"])
        assert f"# n_uniques {exp['n_uniques']}" in out.stdout
        # the run log in the reference's BenchmarkSingleton format
        # (cpp/BenchmarkSingleton.hpp:104-115, messages of PSDBSCAN.hcu:986-997,1113,1198,1279-1292)
        lines = runlog.read_text().split("\n")
        assert lines[0] == "Timer/Message\tValue"
        log_rows = [ln.split("\t") for ln in lines[1:] if ln]
        msgs = [b for a, b in log_rows if a.startswith("At ") and a.endswith(" sec")]
        timers = {a: b for a, b in log_rows if not a.startswith("At ")}
        assert msgs[0] == f"DBSCAN settings: epsilon = {eps:.6f}" and msgs[1] == f"DBSCAN settings: minPts = {minpts}"
        assert ("DBSCAN settings: using Hamming distance measure instead of Levenshtein distance" in msgs) == bool(ham)
        assert f"Total number of sequences = {len(got)}" in msgs
        assert f"Number of uniques = {exp['n_uniques']}" in msgs
        assert f"Number of core points = {exp['n_cores']}" in msgs
        assert f"Total number of clusters = {int(exp['clone_id'].max())}" in msgs
        n_noise = int((exp["clone_id"] == 0).sum())
        assert f"Number of noise points = {n_noise}" in msgs
        assert f"Percentage of noise points = {100 * n_noise // len(got)}%" in msgs
        depths = [int(m.split(": ")[1]) for m in msgs if m.startswith("Number of trees with depth ")]
        assert len(depths) == 51 and sum(depths) == len(got)
        for name in _lib.STAGE_NAMES:
            assert timers[name].endswith(" sec") and float(timers[name][:-4]) >= 0.0


def test_large_input_paths_forced_on_small_inputs(gpu_ctx):
    """The library radix sort of the entry order (otherwise above 2^18
    entries) and 1,024-pair chunks (otherwise for pair lists of 2^24 or more),
    switched on for inputs the oracle checks in seconds."""
    F = _lib.IDB_FLAG_FORCE_LARGE_PATHS
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    check(rep, configs.C2, fast=True, tag="forced c2", flags=F)
    check(synth.generate(configs.c1_params()), configs.C1, tag="forced c1", flags=F)
    check(synth.generate(configs.c5_params(n_rows=20_000, n_seeds=200)), configs.C5, fast=True, tag="forced c5",
          flags=F)
    # a capacity of four 1,024-pair chunks: the waves overflow it and the
    # fuse pass grows the clusters
    got, _ = check(rep, configs.C2, fast=True, tag="forced c2 overflow", flags=F, edge_capacity=4096)
    assert got.stats["fuse_from_edges"] == 0
    reps = configs.c3_repertoires(n_reps=24, scale=0.05)
    cat = synth.concat(reps)
    got = engine.dbscan_batch(cat, configs.C3, flags=F | _lib.IDB_FLAG_GRAPH)
    for r, rp in enumerate(reps):
        s, e = int(cat.rep_start[r]), int(cat.rep_start[r + 1])
        exp = oracle_of(rp, configs.C3)
        for k in ("clone_id", "core", "copies", "uniques", "counts"):
            assert np.array_equal(getattr(got, k)[s:e], exp[k]), (r, k)


def test_split_tiles_equal_the_oracle(gpu_ctx):
    """Off-diagonal tiles of more than 32 queries run as two tiles
    (IDB_FLAG_SPLIT_TILES): the same results on the one-word, multi-word,
    Hamming and batched paths, with more tiles than the plain run."""
    F = _lib.IDB_FLAG_SPLIT_TILES
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    got, _ = check(rep, configs.C2, fast=True, tag="split c2", flags=F)
    assert got.stats["tiles"] > engine.dbscan(rep, configs.C2).stats["tiles"]
    check(synth.generate(configs.c1_params()), configs.C1, tag="split c1", flags=F)
    rep5 = synth.generate(configs.c5_params(n_rows=20_000, n_seeds=200))
    got, _ = check(rep5, configs.C5, fast=True, tag="split c5", flags=F)
    assert got.stats["tiles"] > engine.dbscan(rep5, configs.C5).stats["tiles"]
    reps = configs.c3_repertoires(n_reps=24, scale=0.05)
    cat = synth.concat(reps)
    got = engine.dbscan_batch(cat, configs.C3, flags=F | _lib.IDB_FLAG_GRAPH)
    for r, rp in enumerate(reps):
        s, e = int(cat.rep_start[r]), int(cat.rep_start[r + 1])
        exp = oracle_of(rp, configs.C3)
        for k in ("clone_id", "core", "copies", "uniques", "counts"):
            assert np.array_equal(getattr(got, k)[s:e], exp[k]), (r, k)


def test_tree_depth_histogram_measures_the_forest_before_compression(gpu_ctx):
    """The depth histogram (the reference's tree_depth_evaluation_kernel,
    PSDBSCAN.hcu:154-177, run right after cluster growth) on a forest whose
    shape is fixed: row 0 (100 letters) is a star centre one substitution from
    100 distinct variants (two apart from each other), so every variant is a core row
    linked straight under row 0, and 5 duplicates of the first variant are
    two steps down; from the pair list and from the fuse pass after a
    pair-list overflow (100 pairs, capacity 64) alike."""
    variants = ["A" * p + "C" + "A" * (99 - p) for p in range(100)]
    rows = [("A" * 100, [3], [1])] + [(v, [3], [1]) for v in variants] + [(variants[0], [3], [1])] * 5
    rep = rep_from_rows(rows)
    cl = configs.Clustering(eps=0.0, abs_k=1, min_pts=2)
    for cap in (0, 1):
        got = engine.dbscan(rep, cl, tree_stats=True, edge_capacity=cap)
        assert got.stats["fuse_from_edges"] == (1 if cap == 0 else 0)
        hist = got.stats["tree_depth"]
        assert list(hist[:4]) == [1, 100, 5, 0], (cap, list(hist[:6]))
        assert sum(hist) == rep.n
