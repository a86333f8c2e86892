"""The learned candidate filter: encoder handle, training-pair assembly,
threshold tuning and the learned clustering run, over the C ABI.

Pairs are labelled by the exact device kernels (idb_neighbour_pairs for
positives, idb_pair_distances for sampled negatives); the encoder trains on
the device (idb_encoder_train); idb_dbscan_learned proposes pairs by embedding
dot product and verifies each proposal exactly.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .configs import Clustering
from .engine import Result, _dataset, _params, default_context
from .synth import Repertoire

D = 64
P = 32


def vocab_for(alphabet: bytes) -> np.ndarray:
    """Token of every byte value: the alphabet's letters get 0 .. n-1, every
    other byte token 31."""
    v = np.full(256, 31, np.uint8)
    for i, b in enumerate(alphabet[:31]):
        v[b] = i
    return v


@dataclass
class TrainParams:
    epochs: int = 2
    batch_pairs: int = 4096
    lr: float = 2e-3
    margin: float = 0.9
    seed: int = 1
    log_every: int = 0


class Encoder:
    """An idb_encoder: parameters and optimiser state on one device."""

    def __init__(self, vocab: np.ndarray, seed: int = 0, device: int = 0, positions: int = P):
        """positions: letters per row the encoder reads, 32 (protein rows) or
        160 (nucleotide rows of up to 160 letters)."""
        self.vocab = np.ascontiguousarray(vocab, np.uint8)
        h = C.c_void_p()
        rc = _lib.lib.idb_encoder_create_positions(device, self.vocab.ctypes.data, positions, seed, C.byref(h))
        if rc != _lib.IDB_OK:
            raise _lib.IdbError(rc, "idb_encoder_create_positions failed")
        self._h = h
        self.device = device
        self.positions = positions

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib.idb_encoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def check(self, rc: int) -> None:
        if rc != _lib.IDB_OK:
            raise _lib.IdbError(rc, _lib.lib.idb_encoder_last_error(self._h).decode(errors="replace"))

    def get_params(self) -> np.ndarray:
        out = np.zeros(_lib.IDB_ENC_PARAMS, np.float32)
        self.check(_lib.lib.idb_encoder_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == _lib.IDB_ENC_PARAMS
        self.check(_lib.lib.idb_encoder_set_params(self._h, flat.ctypes.data))

    def embed(self, rep: Repertoire) -> np.ndarray:
        z = np.zeros((rep.n, D), np.float32)
        self.check(_lib.lib.idb_encoder_embed(self._h, _dataset(rep), z.ctypes.data))
        return z

    def loss_grad(self, rep: Repertoire, pairs: np.ndarray, labels: np.ndarray, margin: float):
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        labels = np.ascontiguousarray(labels, np.float32)
        loss = C.c_float()
        g = np.zeros(_lib.IDB_ENC_PARAMS, np.float32)
        self.check(_lib.lib.idb_encoder_loss_grad(self._h, _dataset(rep), pairs.ctypes.data, labels.ctypes.data,
                                                  len(labels), margin, C.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def train(self, rep: Repertoire, pairs: np.ndarray, labels: np.ndarray, tp: TrainParams) -> dict:
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        labels = np.ascontiguousarray(labels, np.float32)
        p = _lib.IdbTrainParams(tp.epochs, tp.batch_pairs, tp.lr, tp.margin, tp.seed, tp.log_every, 0)
        st = _lib.IdbTrainStats()
        self.check(_lib.lib.idb_encoder_train(self._h, _dataset(rep), pairs.ctypes.data, labels.ctypes.data,
                                              len(labels), C.byref(p), C.byref(st)))
        return st.as_dict()


def pair_distances(rep: Repertoire, pairs: np.ndarray, hamming: bool = False, ctx=None) -> np.ndarray:
    ctx = ctx or default_context()
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    out = np.zeros(len(pairs), np.uint32)
    ctx.check(_lib.lib.idb_pair_distances(ctx.handle, _dataset(rep), pairs.ctypes.data, len(pairs),
                                          1 if hamming else 0, out.ctypes.data))
    return out


def neighbour_pairs(rep: Repertoire, cl: Clustering, ctx=None, cap: int | None = None) -> np.ndarray:
    """Unordered unique-row pairs (x < y) with N(x, y), from the exact count
    pass; the pass's own pair list is given the same capacity (dense clones,
    as in the nucleotide workloads, hold tens of pairs per row)."""
    ctx = ctx or default_context()
    cap = cap if cap is not None else min(max(1 << 22, 256 * rep.n), 1 << 28)
    out = np.zeros((cap, 2), np.uint32)
    n = C.c_uint64()
    ctx.check(_lib.lib.idb_neighbour_pairs(ctx.handle, _dataset(rep), _params(cl, False, 0, edge_capacity=cap),
                                           out.ctypes.data, cap, C.byref(n)))
    return out[: n.value].copy()


def dbscan_learned(rep: Repertoire, cl: Clustering, enc: Encoder, min_sim: float, ctx=None,
                   counts: bool = True, edge_capacity: int = 0) -> Result:
    ctx = ctx or default_context()
    n = rep.n
    cid = np.zeros(n, np.uint32)
    core = np.zeros(n, np.int32)
    cop = np.zeros(n, np.uint32)
    unq = np.zeros(n, np.uint32)
    cnt = np.zeros(n, np.uint32) if counts else None
    outs = _lib.make_outputs(cid, core, cop, unq, cnt)
    st = _lib.IdbStats()
    ctx.check(_lib.lib.idb_dbscan_learned(ctx.handle, enc.handle, float(min_sim), _dataset(rep),
                                          _params(cl, False, 0, edge_capacity), outs, st))
    return Result(cid, core, cop, unq, int(st.n_uniques), cnt, st.as_dict())


def gene_group_keys(rep: Repertoire) -> np.ndarray:
    """Lowest V gene and lowest J gene of every row as one key (-1: none)."""
    def lowest(words: np.ndarray) -> np.ndarray:
        w = words.reshape(rep.n, -1).astype(np.uint32)
        out = np.full(rep.n, -1, np.int64)
        for k in range(w.shape[1] - 1, -1, -1):
            col = w[:, k]
            nz = col != 0
            low = (col & (~col + np.uint32(1))).astype(np.uint64)
            out[nz] = k * 32 + np.log2(low[nz]).astype(np.int64)
        return out

    v, j = lowest(rep.vbits), lowest(rep.jbits)
    return np.where((v >= 0) & (j >= 0), v * 4096 + j, -1)


def sample_negatives(rep: Repertoire, n: int, max_dlen: int, seed: int, window: int = 64) -> np.ndarray:
    """Random pairs of rows that share their lowest V and J gene and differ in
    length by at most max_dlen: the pairs the filter has to reject."""
    rng = np.random.default_rng(seed)
    key = gene_group_keys(rep)
    ok = np.nonzero(key >= 0)[0]
    order = ok[np.lexsort((rep.length[ok], key[ok]))]
    m = len(order)
    if m < 2:
        return np.zeros((0, 2), np.uint32)
    i = rng.integers(0, m - 1, size=4 * n)
    j = np.minimum(i + rng.integers(1, window + 1, size=4 * n), m - 1)
    a, b = order[i], order[j]
    keep = (a != b) & (key[a] == key[b]) & (np.abs(rep.length[a].astype(np.int64) - rep.length[b]) <= max_dlen)
    pr = np.stack([a[keep], b[keep]], axis=1).astype(np.uint32)
    return pr[:n]


def training_pairs(rep: Repertoire, cl: Clustering, n_pairs: int, seed: int = 0, ctx=None):
    """n_pairs labelled pairs, half of them neighbours (label 1) from the
    exact count pass and half sampled non-neighbours (label 0) whose exact
    distance exceeds the threshold."""
    rng = np.random.default_rng(seed)
    pos = neighbour_pairs(rep, cl, ctx)
    n_pos = min(n_pairs // 2, len(pos))
    pos = pos[rng.choice(len(pos), n_pos, replace=False)] if n_pos < len(pos) else pos
    t = cl.abs_k if cl.abs_k >= 0 else int(np.float32(rep.length.max()) * np.float32(cl.eps))
    neg = sample_negatives(rep, 2 * (n_pairs - n_pos), t, seed + 1)
    if len(neg):
        d = pair_distances(rep, neg, ctx=ctx)
        neg = neg[d > t][: n_pairs - n_pos]
    pairs = np.concatenate([pos, neg]).astype(np.uint32)
    labels = np.concatenate([np.ones(len(pos), np.float32), np.zeros(len(neg), np.float32)])
    perm = rng.permutation(len(labels))
    return pairs[perm], labels[perm]


def similarity(z: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    return (z[pairs[:, 0]] * z[pairs[:, 1]]).sum(axis=1)


def tune_threshold(z: np.ndarray, pos_pairs: np.ndarray, recall: float = 0.999, slack: float = 0.0) -> float:
    """The largest threshold that keeps `recall` of the given neighbour pairs
    (embedding dot products are compared in bfloat16 inputs on the device, so
    a small slack absorbs rounding)."""
    s = np.sort(similarity(z, pos_pairs))
    k = int(np.floor((1.0 - recall) * len(s)))
    return float(s[min(k, len(s) - 1)] - slack)


def ari(a: np.ndarray, b: np.ndarray) -> float:
    """Adjusted Rand index of two labelings; noise (0) rows are singletons."""
    from sklearn.metrics import adjusted_rand_score

    def canon(x):
        x = np.asarray(x, np.int64).copy()
        noise = x == 0
        x[noise] = -1 - np.arange(noise.sum())
        return x

    return float(adjusted_rand_score(canon(a), canon(b)))
