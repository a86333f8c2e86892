"""ctypes wrapper of liboracle.so, the CPU restatement of the reference's
clustering semantics.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the timed CPU
reference, never as a compute path of the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class OrcData(C.Structure):
    _fields_ = [("n", C.c_uint32), ("seq", C.c_void_p), ("off", C.c_void_p), ("len", C.c_void_p),
                ("vbits", C.c_void_p), ("vw", C.c_uint32), ("jbits", C.c_void_p), ("jw", C.c_uint32)]


class OrcParams(C.Structure):
    _fields_ = [("eps", C.c_float), ("abs_k", C.c_int32), ("min_pts", C.c_uint32), ("use_hamming", C.c_int32),
                ("nthreads", C.c_int32)]


class OrcResult(C.Structure):
    _fields_ = [("clone_id", C.c_void_p), ("core", C.c_void_p), ("copies", C.c_void_p), ("uniques", C.c_void_p),
                ("counts", C.c_void_p), ("n_uniques", C.c_uint32), ("n_cores", C.c_uint32),
                ("n_clusters", C.c_uint32), ("n_noise", C.c_uint32), ("n_edges", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.orc_threshold.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_int32]
        L.orc_threshold.restype = C.c_int32
        L.orc_levenshtein.argtypes = [C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32]
        L.orc_levenshtein.restype = C.c_uint32
        L.orc_hamming.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32]
        L.orc_hamming.restype = C.c_uint32
        L.orc_lev_within.argtypes = [C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_int32]
        L.orc_lev_within.restype = C.c_int32
        L.orc_gene_bitsets.argtypes = [C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, C.c_uint32]
        L.orc_gene_bitsets.restype = None
        for f in ("orc_neighbour_count_pass", "orc_neighbour_fuse_pass"):
            getattr(L, f).argtypes = [C.POINTER(OrcData), C.POINTER(OrcParams), C.c_uint32, C.c_uint32]
            getattr(L, f).restype = C.c_int32
        for f in ("orc_ref_dbscan", "orc_fast_dbscan"):
            getattr(L, f).argtypes = [C.POINTER(OrcData), C.POINTER(OrcParams), C.POINTER(OrcResult)]
            getattr(L, f).restype = C.c_int32
        L.orc_sample_rows.argtypes = [C.POINTER(OrcData), C.POINTER(OrcParams), vp, C.c_uint32, vp, vp, vp, vp, vp,
                                      C.c_uint64]
        L.orc_sample_rows.restype = C.c_int32
        _lib = L
    return _lib


def threshold(lx: int, ly: int, eps: float, abs_k: int = -1) -> int:
    return lib().orc_threshold(lx, ly, C.c_float(eps), abs_k)


def levenshtein(a: bytes, b: bytes) -> int:
    return lib().orc_levenshtein(a, len(a), b, len(b))


def hamming(a: bytes, b: bytes) -> int:
    assert len(a) == len(b)
    return lib().orc_hamming(a, b, len(a))


def lev_within(a: bytes, b: bytes, t: int) -> bool:
    return bool(lib().orc_lev_within(a, len(a), b, len(b), t))


def gene_bitsets(calls: list[bytes], genes: list[bytes]) -> np.ndarray:
    width = (len(genes) + 31) // 32
    out = np.zeros((len(calls), max(width, 1)), np.int32)

    def pack(ss):
        lens = np.array([len(s) for s in ss], np.uint32)
        off = np.zeros(len(ss), np.uint64)
        if len(ss):
            off[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
        data = np.frombuffer(b"".join(ss) + b"\0", np.uint8).copy()
        return data, off, lens

    cd, co, cl = pack(calls)
    gd, go, gl = pack(genes)
    lib().orc_gene_bitsets(len(calls), cd.ctypes.data, co.ctypes.data, cl.ctypes.data, len(genes), gd.ctypes.data,
                           go.ctypes.data, gl.ctypes.data, out.ctypes.data, max(width, 1))
    return out[:, :width]


class Data:
    """Keeps the arrays alive for an OrcData view."""

    def __init__(self, seq, off, length, vbits, jbits):
        self.seq = np.ascontiguousarray(seq if len(seq) else np.zeros(1, np.uint8), dtype=np.uint8)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.len = np.ascontiguousarray(length, dtype=np.uint32)
        self.vb = np.ascontiguousarray(vbits, dtype=np.int32)
        self.jb = np.ascontiguousarray(jbits, dtype=np.int32)
        d = OrcData()
        d.n = len(self.len)
        d.seq, d.off, d.len = self.seq.ctypes.data, self.off.ctypes.data, self.len.ctypes.data
        d.vbits, d.vw = self.vb.ctypes.data, self.vb.shape[1] if self.vb.ndim == 2 else 0
        d.jbits, d.jw = self.jb.ctypes.data, self.jb.shape[1] if self.jb.ndim == 2 else 0
        self.c = d

    @classmethod
    def of(cls, rep) -> "Data":
        return cls(rep.seq, rep.off, rep.length, rep.vbits, rep.jbits)


def params(eps=0.15, abs_k=-1, min_pts=2, use_hamming=False, nthreads=0) -> OrcParams:
    p = OrcParams()
    p.eps = float(np.float32(eps))
    p.abs_k = int(abs_k)
    p.min_pts = int(min_pts) & 0xFFFFFFFF
    p.use_hamming = 1 if use_hamming else 0
    p.nthreads = int(nthreads)
    return p


def dbscan(data: Data, prm: OrcParams, fast: bool = False) -> dict:
    n = data.c.n
    out = {k: np.zeros(n, t) for k, t in (("clone_id", np.uint32), ("core", np.int32), ("copies", np.uint32),
                                          ("uniques", np.uint32), ("counts", np.uint32))}
    r = OrcResult()
    r.clone_id, r.core, r.copies = out["clone_id"].ctypes.data, out["core"].ctypes.data, out["copies"].ctypes.data
    r.uniques, r.counts = out["uniques"].ctypes.data, out["counts"].ctypes.data
    f = lib().orc_fast_dbscan if fast else lib().orc_ref_dbscan
    rc = f(C.byref(data.c), C.byref(prm), C.byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    out.update(n_uniques=r.n_uniques, n_cores=r.n_cores, n_clusters=r.n_clusters, n_noise=r.n_noise,
               n_edges=r.n_edges)
    return out


def neighbour(data: Data, prm: OrcParams, x: int, y: int, fuse: bool = False) -> bool:
    f = lib().orc_neighbour_fuse_pass if fuse else lib().orc_neighbour_count_pass
    return bool(f(C.byref(data.c), C.byref(prm), x, y))


def sample_rows(data: Data, prm: OrcParams, rows, nbr_cap: int = 1 << 26) -> dict:
    """Representative, class size, exact count and count-pass neighbour rows
    of each sampled row (orc_sample_rows), for inputs too large to cluster."""
    xs = np.ascontiguousarray(rows, np.uint32)
    m = len(xs)
    rep, cop, cnt = (np.zeros(m, np.uint32) for _ in range(3))
    off = np.zeros(m + 1, np.uint64)
    nbr = np.zeros(max(nbr_cap, 1), np.uint32)
    rc = lib().orc_sample_rows(C.byref(data.c), C.byref(prm), xs.ctypes.data, m, rep.ctypes.data, cop.ctypes.data,
                               cnt.ctypes.data, off.ctypes.data, nbr.ctypes.data, nbr_cap)
    if rc != 0:
        raise RuntimeError(f"oracle sample failed with {rc}")
    total = int(off[m])
    if total > nbr_cap:
        raise RuntimeError(f"{total} neighbour rows exceed the buffer of {nbr_cap}")
    return {"rep": rep, "copies": cop, "count": cnt,
            "neighbours": [nbr[int(off[s]):int(off[s + 1])] for s in range(m)]}
