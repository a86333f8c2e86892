"""The learned candidate filter.

CPU: the encoder oracle is pinned by independent definitions — bfloat16
rounding against torch's conversion, GELU against torch's tanh GELU, the
convolution against scipy's correlate1d, the backward pass against central
finite differences in float64.

GPU: the device encoder's forward pass and gradients against the oracle (with
the same bfloat16 rounding points), the learned clustering run against the
exact run when every gated pair is proposed (bit-exact), and a trained
encoder's recall / ARI and the candidate reduction it buys over an untrained
one and over the exact filter.
"""
from __future__ import annotations

import numpy as np
import pytest

from immunodbscan_amd import _lib, configs, synth
from oracle import learned as Lr


def small_rep(n=4000, seeds=400, seed=21):
    return synth.generate(configs.c2_params(n_rows=n, n_seeds=seeds, seed=seed))


# --------------------------------------------------------------------------
# CPU: oracle pins
def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(0)
    x = np.concatenate([(rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000)).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.3895314e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(Lr.bf16(x), ref)


def test_gelu_matches_torch_tanh_gelu():
    import torch

    x = np.linspace(-8, 8, 4001)
    ref = torch.nn.functional.gelu(torch.from_numpy(x), approximate="tanh").numpy()
    assert np.allclose(Lr.gelu(x), ref, atol=1e-12)
    # derivative against central differences
    h = 1e-6
    assert np.allclose(Lr.gelu_grad(x), (Lr.gelu(x + h) - Lr.gelu(x - h)) / (2 * h), atol=1e-7)


def test_conv_matches_scipy_correlate():
    from scipy.ndimage import correlate1d

    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, Lr.P, 8))
    Wt = rng.standard_normal((5, Lr.T * 8))
    b = rng.standard_normal(5)
    H = Lr.conv(X, Wt, b)
    W = Wt.reshape(5, Lr.T, 8)
    for o in range(5):
        ref = sum(correlate1d(X[:, :, c], W[o, :, c], axis=1, mode="constant", origin=0) for c in range(8)) + b[o]
        assert np.allclose(H[:, :, o], ref, atol=1e-12)


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    flat = Lr.init_params(3).astype(np.float64)
    tok = rng.integers(0, 20, (6, Lr.P))
    lens = np.array([8, 13, 25, 32, 1, 0])
    ia, ib, y = np.array([0, 1, 2, 3]), np.array([3, 4, 5, 1]), np.array([1.0, 0.0, 1.0, 0.0])
    _, g, _ = Lr.loss_and_grad(flat, tok, lens, ia, ib, y, 0.9, emulate_bf16=False, dtype=np.float64)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in Lr.SHAPES])
    # a few coordinates of every tensor (embedding rows of used tokens)
    idx = [int(tok[0, 0]) * Lr.C0 + 3]
    for k in range(1, len(Lr.SHAPES)):
        idx += list(rng.integers(offs[k], offs[k + 1], 6))
    h = 1e-6
    for i in idx:
        f = flat.copy()
        f[i] += h
        lp = Lr.loss_and_grad(f, tok, lens, ia, ib, y, 0.9, False, np.float64)[0]
        f[i] -= 2 * h
        lm = Lr.loss_and_grad(f, tok, lens, ia, ib, y, 0.9, False, np.float64)[0]
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_vocab_and_tokens():
    from immunodbscan_amd import learned

    v = learned.vocab_for(b"ACDEFGHIKLMNPQRSTVWY")
    assert v[ord("A")] == 0 and v[ord("Y")] == 19 and v[ord("Z")] == 31 and v.max() == 31
    tok, lens = Lr.tokens_of([b"ACD", b"Y" * 40, b""], v)
    assert list(lens) == [3, 32, 0] and list(tok[0, :4]) == [0, 1, 2, 0] and tok[1, 31] == 19


# --------------------------------------------------------------------------
# GPU
def _vocab():
    from immunodbscan_amd import learned

    return learned.vocab_for(synth.LETTERS[synth.AA])


def _tokens(rep, vocab):
    seqs = [bytes(rep.seq[rep.off[i]:rep.off[i] + rep.length[i]]) for i in range(rep.n)]
    return Lr.tokens_of(seqs, vocab)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_encoder_forward_matches_oracle(gpu_ctx):
    from immunodbscan_amd import learned

    rep = small_rep().subset(np.arange(96))
    enc = learned.Encoder(_vocab(), seed=5)
    flat = enc.get_params()
    z = enc.embed(rep)
    tok, lens = _tokens(rep, enc.vocab)
    ref = Lr.forward(flat, tok, lens, emulate_bf16=True)["z"]
    print("forward max abs diff", np.abs(z - ref).max(), flush=True)
    assert np.abs(z - ref).max() < 1e-2 and np.abs(z - ref).mean() < 1e-3
    assert np.allclose(np.linalg.norm(z, axis=1), 1.0, atol=1e-5)


@pytest.mark.gpu
def test_encoder_gradients_match_oracle(gpu_ctx):
    from immunodbscan_amd import learned

    rep = small_rep().subset(np.arange(128))
    enc = learned.Encoder(_vocab(), seed=6)
    flat = enc.get_params()
    rng = np.random.default_rng(7)
    pairs = rng.integers(0, rep.n, (48, 2)).astype(np.uint32)
    labels = (rng.random(48) < 0.5).astype(np.float32)
    loss, g = enc.loss_grad(rep, pairs, labels, 0.9)
    tok, lens = _tokens(rep, enc.vocab)
    lref, gref, _ = Lr.loss_and_grad(flat, tok, lens, pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64),
                                     labels, 0.9, emulate_bf16=True)
    print("loss", loss, lref, flush=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref) + 1e-6
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in Lr.SHAPES])
    for k, (name, _) in enumerate(Lr.SHAPES):
        a, b = g[offs[k]:offs[k + 1]], gref[offs[k]:offs[k + 1]]
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(name, "rel", _rel(a, b), "cos", cos, flush=True)
        assert _rel(a, b) < 2e-2 and cos > 0.999, name


def _nt_rep(n, seeds, seed):
    """Nucleotide rows of 24-150 letters, as the 20M-row workload draws them."""
    return synth.generate(configs.c5_params(n_rows=n, n_seeds=seeds, seed=seed))


@pytest.mark.gpu
def test_encoder_160_positions_forward_and_gradients_match_oracle(gpu_ctx):
    """The nucleotide encoder: 160 positions per row (rows of 24-150 letters)."""
    from immunodbscan_amd import learned

    rep = _nt_rep(64, 64, 51)
    enc = learned.Encoder(learned.vocab_for(synth.LETTERS[synth.NT]), seed=5, positions=160)
    assert _lib.lib.idb_encoder_positions(enc._h) == 160
    flat = enc.get_params()
    z = enc.embed(rep)
    seqs = [bytes(rep.seq[rep.off[i]:rep.off[i] + rep.length[i]]) for i in range(rep.n)]
    tok, lens = Lr.tokens_of(seqs, enc.vocab, positions=160)
    assert tok.shape == (64, 160) and lens.max() > 96
    ref = Lr.forward(flat, tok, lens, emulate_bf16=True)["z"]
    print("forward max abs diff", np.abs(z - ref).max(), flush=True)
    assert np.abs(z - ref).max() < 1e-2 and np.abs(z - ref).mean() < 1e-3
    rng = np.random.default_rng(8)
    pairs = rng.integers(0, rep.n, (24, 2)).astype(np.uint32)
    labels = (rng.random(24) < 0.5).astype(np.float32)
    loss, g = enc.loss_grad(rep, pairs, labels, 0.9)
    lref, gref, _ = Lr.loss_and_grad(flat, tok, lens, pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64),
                                     labels, 0.9, emulate_bf16=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref) + 1e-6
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in Lr.SHAPES])
    for k, (name, _) in enumerate(Lr.SHAPES):
        a, b = g[offs[k]:offs[k + 1]], gref[offs[k]:offs[k + 1]]
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        assert _rel(a, b) < 2e-2 and cos > 0.999, name
    # rows longer than the encoder reads are refused by a learned run
    short = learned.Encoder(learned.vocab_for(synth.LETTERS[synth.NT]), seed=5)
    with pytest.raises(_lib.IdbError) as e:
        learned.dbscan_learned(rep, configs.C5, short, -2.0)
    assert e.value.code == _lib.IDB_ERR_UNSUPPORTED


@pytest.mark.gpu
def test_learned_run_on_long_rows_proposing_everything_equals_exact(gpu_ctx):
    """Rows of up to 5 pattern words: every gated pair proposed, verified by
    the multi-word (banded) Myers batches, gives the exact run's results."""
    from immunodbscan_amd import engine, learned

    rep = _nt_rep(6_000, 300, 52)
    enc = learned.Encoder(learned.vocab_for(synth.LETTERS[synth.NT]), seed=1, positions=160)
    ex = engine.dbscan(rep, configs.C5)
    le = learned.dbscan_learned(rep, configs.C5, enc, -2.0)
    for f in ("clone_id", "core", "copies", "counts"):
        assert np.array_equal(getattr(ex, f), getattr(le, f)), f
    assert le.stats["neighbour_pairs"] == ex.stats["neighbour_pairs"] > 0
    assert le.stats["gated_pairs"] == ex.stats["gated_pairs"]
    # the same exact 3-gram bound before Myers
    assert le.stats["myers_pairs"] == ex.stats["myers_pairs"]


@pytest.mark.gpu
def test_learned_run_proposing_everything_equals_exact(gpu_ctx):
    from immunodbscan_amd import engine, learned

    rep = small_rep(20_000, 2_000, seed=22)
    enc = learned.Encoder(_vocab(), seed=1)
    ex = engine.dbscan(rep, configs.C2)
    le = learned.dbscan_learned(rep, configs.C2, enc, -2.0)
    for f in ("clone_id", "core", "copies", "counts"):
        assert np.array_equal(getattr(ex, f), getattr(le, f)), f
    assert le.stats["neighbour_pairs"] == ex.stats["neighbour_pairs"]
    # every gated pair proposed; the same exact letter bound before Myers
    assert le.stats["gated_pairs"] == ex.stats["gated_pairs"]
    assert le.stats["myers_pairs"] == ex.stats["myers_pairs"]
    # a pair list far too short: the learned pass runs again with a grown
    # list, with the same results and statistics
    sm = learned.dbscan_learned(rep, configs.C2, enc, -2.0, edge_capacity=64)
    assert sm.stats["fuse_from_edges"] == 1
    for f in ("clone_id", "core", "copies", "counts"):
        assert np.array_equal(getattr(sm, f), getattr(le, f)), f
    for k in ("gated_pairs", "neighbour_pairs", "myers_pairs"):
        assert sm.stats[k] == le.stats[k], k


@pytest.mark.gpu
def test_trained_filter_recall_ari_and_gain(gpu_ctx):
    from immunodbscan_amd import engine, learned

    cl = configs.C2
    train = small_rep(100_000, 10_000, seed=31)
    pairs, labels = learned.training_pairs(train, cl, 600_000, seed=3)
    assert 0.4 < labels.mean() < 0.6
    enc = learned.Encoder(_vocab(), seed=2)
    base = learned.Encoder(_vocab(), seed=2)  # the same network, untrained
    st = enc.train(train, pairs, labels, learned.TrainParams(epochs=6, batch_pairs=4096, lr=5e-4, margin=0.9))
    print("train", st, flush=True)
    assert st["last_loss"] < 0.8 * st["first_loss"]

    test = small_rep(40_000, 4_000, seed=32)
    tune = small_rep(40_000, 4_000, seed=33)
    pos_tune = learned.neighbour_pairs(tune, cl)
    tau = learned.tune_threshold(enc.embed(tune), pos_tune, recall=0.999, slack=0.01)
    tau0 = learned.tune_threshold(base.embed(tune), pos_tune, recall=0.999, slack=0.01)
    ex = engine.dbscan(test, cl)
    le = learned.dbscan_learned(test, cl, enc, tau)
    lb = learned.dbscan_learned(test, cl, base, tau0)
    recall = le.stats["neighbour_pairs"] / ex.stats["neighbour_pairs"]
    a = learned.ari(ex.clone_id, le.clone_id)
    print(f"tau {tau:.4f} recall {recall:.5f} ari {a:.5f} proposals trained {le.stats['gated_pairs']} "
          f"untrained {lb.stats['gated_pairs']} exact gated {ex.stats['gated_pairs']} "
          f"Myers runs learned {le.stats['myers_pairs']} exact {ex.stats['myers_pairs']}", flush=True)
    assert recall >= 0.995 and a >= 0.995
    # what training buys: far fewer proposals than the untrained network at
    # the same recall target, and than the exact path's gated pairs
    assert le.stats["gated_pairs"] * 3 < lb.stats["gated_pairs"]
    assert le.stats["gated_pairs"] * 5 < ex.stats["gated_pairs"]
    # and never more Myers runs than the exact path's own letter bound leaves
    assert le.stats["myers_pairs"] <= ex.stats["myers_pairs"]
