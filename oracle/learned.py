"""CPU restatement of the learned candidate filter's encoder.  TEST INFRASTRUCTURE
ONLY: the tests use it to check the GPU encoder's forward pass and gradients.

The encoder has no counterpart in the reference; it is the small 1-D CNN over
CDR3 letters that the learned filter uses to propose neighbour pairs:

    tokens -> embedding (32 tokens x 32)
           -> 3 x [conv1d(kernel 5, 'same' zero padding, 128 channels) + GELU]
           -> mean over the row's letters -> linear 128 -> 64 -> L2 normalise.

A row occupies P positions (32 by default; the token matrix's width sets
it); positions at or past its length hold zeros
after every layer, so each row is convolved with its own zero padding.  GELU
is the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).

With ``emulate_bf16`` the restatement rounds where the GPU stores bfloat16
(round to nearest even): the weights, the embedding rows, every activation
after GELU, the saved pre-activations and the pre-activation gradients of
the convolution layers; everything else is float32 (or float64 when asked).

Flat parameter order (shared with the GPU buffer):
    emb [32, 32], w1 [128, 5*32], b1 [128], w2 [128, 5*128], b2 [128],
    w3 [128, 5*128], b3 [128], wp [64, 128], bp [64]
Convolution weights are stored out-channel major, k = tap * C_in + c_in.

Contrastive pair loss over B pairs (a, b, y):  d = |z_a - z_b|,
    L = (1/B) sum [ y d^2 + (1 - y) max(0, m - d)^2 ].
"""
from __future__ import annotations

import numpy as np

P, T, NTOK, C0, C, D = 32, 5, 32, 32, 128, 64
SHAPES = [("emb", (NTOK, C0)), ("w1", (C, T * C0)), ("b1", (C,)), ("w2", (C, T * C)), ("b2", (C,)),
          ("w3", (C, T * C)), ("b3", (C,)), ("wp", (D, C)), ("bp", (D,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
GELU_K = np.sqrt(2.0 / np.pi)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unflatten(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n]).reshape(shape)
        o += n
    return out


def flatten(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def init_params(seed: int) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, N(0, 1)
    embedding (the initialisation the GPU encoder uses is its own; this one
    only feeds the CPU tests)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name == "emb":
            d[name] = rng.standard_normal(shape).astype(np.float32)
        elif name.startswith("b"):
            d[name] = np.zeros(shape, np.float32)
        else:
            a = 1.0 / np.sqrt(shape[1])
            d[name] = rng.uniform(-a, a, shape).astype(np.float32)
    return flatten(d)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_K * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = GELU_K * (x + 0.044715 * x ** 3)
    th = np.tanh(u)
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * GELU_K * (1.0 + 3 * 0.044715 * x * x)


def _im2col(X: np.ndarray) -> np.ndarray:
    """[n, P, Cin] -> [n, P, T*Cin] with column t*Cin + c = X[:, p + t - 2, c]."""
    n, npos, cin = X.shape
    Xp = np.zeros((n, npos + T - 1, cin), X.dtype)
    Xp[:, 2:2 + npos] = X
    return np.concatenate([Xp[:, t:t + npos] for t in range(T)], axis=2)


def conv(X: np.ndarray, Wt: np.ndarray, b: np.ndarray) -> np.ndarray:
    return _im2col(X) @ Wt.T + b


def conv_backward_data(dH: np.ndarray, Wt: np.ndarray, cin: int) -> np.ndarray:
    """dX[q, c] = sum_{t, o} dH[q - t + 2, o] Wt[o, t*cin + c]."""
    n, npos = dH.shape[0], dH.shape[1]
    dcol = dH @ Wt  # [n, P, T*cin]: gradient w.r.t. im2col columns
    dX = np.zeros((n, npos + T - 1, cin), dH.dtype)
    for t in range(T):
        dX[:, t:t + npos] += dcol[:, :, t * cin:(t + 1) * cin]
    return dX[:, 2:2 + npos]


def forward(flat: np.ndarray, tok: np.ndarray, lens: np.ndarray, emulate_bf16: bool = True,
            dtype=np.float32) -> dict:
    """tok [n, P] token ids, lens [n] (<= P).  Returns z and the saved
    intermediates."""
    rnd = bf16 if emulate_bf16 else (lambda a: a)
    p = {k: (rnd(v.astype(np.float32)) if emulate_bf16 else v).astype(dtype) for k, v in unflatten(flat).items()}
    n = tok.shape[0]
    lens = np.asarray(lens, np.int64)
    mask = (np.arange(tok.shape[1])[None, :] < lens[:, None]).astype(dtype)[:, :, None]
    X0 = p["emb"][tok] * mask
    cache = {"p": p, "mask": mask, "lens": lens, "tok": tok, "X0": X0, "rnd": rnd}
    A = X0
    for li, (w, b) in enumerate((("w1", "b1"), ("w2", "b2"), ("w3", "b3")), start=1):
        H = conv(A, p[w], p[b]).astype(dtype)
        cache[f"A{li - 1}"] = A
        cache[f"H{li}"] = rnd(H.astype(np.float32)).astype(dtype) if emulate_bf16 else H
        A = (rnd(gelu(H).astype(np.float32)).astype(dtype) if emulate_bf16 else gelu(H)) * mask
    cache["A3"] = A
    L = np.maximum(lens, 1).astype(dtype)[:, None]
    pool = A.sum(axis=1) / L
    pool[lens == 0] = 0
    u = pool @ p["wp"].T + p["bp"]
    nrm = np.maximum(np.sqrt((u * u).sum(axis=1, keepdims=True)), 1e-6)
    z = u / nrm
    cache.update(pool=pool, u=u, nrm=nrm, z=z)
    return cache


def backward(cache: dict, dz: np.ndarray) -> np.ndarray:
    """Gradient of sum(dz * z) with respect to the flat parameters."""
    p, mask, lens, rnd = cache["p"], cache["mask"], cache["lens"], cache["rnd"]
    emulate = rnd is bf16
    z, nrm, pool = cache["z"], cache["nrm"], cache["pool"]
    du = (dz - z * (z * dz).sum(axis=1, keepdims=True)) / nrm
    g = {"wp": du.T @ pool, "bp": du.sum(axis=0)}
    dpool = du @ p["wp"]
    L = np.maximum(lens, 1).astype(dz.dtype)[:, None]
    dpool = np.where(lens[:, None] > 0, dpool / L, 0)
    dA = np.broadcast_to(dpool[:, None, :], cache["A3"].shape) * mask
    for li, (w, b, cin) in zip((3, 2, 1), (("w3", "b3", C), ("w2", "b2", C), ("w1", "b1", C0))):
        dH = dA * gelu_grad(cache[f"H{li}"]) * mask
        if emulate:
            dH = rnd(dH.astype(np.float32)).astype(dz.dtype)
        Aprev = cache[f"A{li - 1}"]
        g[w] = np.einsum("npo,npk->ok", dH, _im2col(Aprev))
        g[b] = dH.sum(axis=(0, 1))
        dA = conv_backward_data(dH, p[w], cin) * mask
    dEmb = np.zeros_like(p["emb"])
    np.add.at(dEmb, cache["tok"].reshape(-1), dA.reshape(-1, C0))
    g["emb"] = dEmb
    return flatten(g)


def pair_loss(z: np.ndarray, ia: np.ndarray, ib: np.ndarray, y: np.ndarray, margin: float):
    """Loss and dL/dz (rows a and b of each pair)."""
    B = len(ia)
    diff = z[ia] - z[ib]
    d2 = (diff * diff).sum(axis=1)
    d = np.sqrt(d2 + 1e-12)
    hinge = np.maximum(0.0, margin - d)
    loss = float((y * d2 + (1 - y) * hinge * hinge).sum() / B)
    coef = (2.0 * y - (1 - y) * 2.0 * hinge / d) / B
    gdiff = coef[:, None] * diff
    dz = np.zeros_like(z)
    np.add.at(dz, ia, gdiff)
    np.add.at(dz, ib, -gdiff)
    return loss, dz


def loss_and_grad(flat, tok, lens, ia, ib, y, margin, emulate_bf16=True, dtype=np.float32):
    cache = forward(flat, tok, lens, emulate_bf16, dtype)
    loss, dz = pair_loss(cache["z"], ia, ib, y, margin)
    return loss, backward(cache, dz.astype(dtype)), cache


def tokens_of(seqs: list[bytes], vocab: np.ndarray, positions: int = P) -> tuple[np.ndarray, np.ndarray]:
    """Token matrix [n, positions] (0 past the end) and lengths (at most
    positions)."""
    tok = np.zeros((len(seqs), positions), np.int64)
    lens = np.zeros(len(seqs), np.int64)
    for i, s in enumerate(seqs):
        b = np.frombuffer(s[:positions], np.uint8)
        tok[i, :len(b)] = vocab[b]
        lens[i] = len(b)
    return tok, lens
