"""Exact pair entry points used to label training data for the learned
filter: idb_pair_distances against the oracle's Wagner-Fischer / Hamming
distances, and idb_neighbour_pairs against a brute-force evaluation of the
reference's neighbour predicate over every pair of unique rows."""
from __future__ import annotations

import random

import numpy as np
import pytest

from immunodbscan_amd import configs, learned, synth
from oracle import oracle as O
from tests.test_gpu_parity import oracle_of, rep_from_rows
from tests.test_oracle import random_rows


def rows_bytes(rep):
    return [bytes(rep.seq[rep.off[i]:rep.off[i] + rep.length[i]]) for i in range(rep.n)]


@pytest.mark.gpu
def test_pair_distances_match_oracle(gpu_ctx):
    rng = random.Random(41)
    rows = []
    for L in list(range(0, 40)) + [63, 64, 65, 96, 128, 160, 200, 255, 256]:
        for _ in range(3):
            alpha = rng.choice([b"ACGT", b"ACDEFGHIKLMNPQRSTVWY", bytes(range(1, 200))])
            rows.append((bytes(rng.choice(alpha) for _ in range(L)), [0], [0]))
    # near neighbours: mutated copies
    for _ in range(60):
        s = bytearray(rng.choice(rows)[0])
        for _ in range(rng.randint(1, 4)):
            if s and rng.random() < 0.5:
                del s[rng.randrange(len(s))]
            else:
                s.insert(rng.randrange(len(s) + 1), rng.choice(b"ACGT"))
        rows.append((bytes(s[:256]), [0], [0]))
    rep = rep_from_rows(rows)
    seqs = rows_bytes(rep)
    pr = np.array([(rng.randrange(rep.n), rng.randrange(rep.n)) for _ in range(3000)], np.uint32)
    d = learned.pair_distances(rep, pr)
    for (a, b), got in zip(pr, d):
        assert got == O.levenshtein(seqs[a], seqs[b]), (a, b)
    h = learned.pair_distances(rep, pr, hamming=True)
    for (a, b), got in zip(pr, h):
        exp = O.hamming(seqs[a], seqs[b]) if len(seqs[a]) == len(seqs[b]) else 0xFFFFFFFF
        assert got == exp, (a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c2-like", "hamming-c1"])
def test_neighbour_pairs_match_brute_force(gpu_ctx, case):
    if case == "c2-like":
        rep = synth.generate(configs.c2_params(n_rows=1200, n_seeds=150, seed=9))
        cl = configs.C2
    else:
        rep = synth.generate(configs.c1_params(n_rows=1200, n_seeds=120, seed=9))
        cl = configs.C1
    got = learned.neighbour_pairs(rep, cl)
    got_set = {(int(a), int(b)) for a, b in got}
    assert len(got_set) == len(got) and all(a < b for a, b in got_set)
    exp = oracle_of(rep, cl)
    uniques = [int(u) for u in exp["uniques"][: exp["n_uniques"]]]
    d = O.Data.of(rep)
    prm = O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming)
    want = {(x, y) for i, x in enumerate(uniques) for y in uniques[i + 1:] if O.neighbour(d, prm, x, y)}
    assert got_set == want
    assert len(got) == exp["n_edges"]
