"""Python mirror of the reference's host->device boundary, InterlinkClass
(cuda/Interlink.hpp:46-69, cuda/Interlink.cu:34-133).

Same constructor arguments in the same order and meaning, same getters; the
work runs in libidb_hip.so.  The constructor performs R2 (gene names found as
substrings of each call string become bitset bits, cuda/BitsetVector.inl:27-45),
packs the sequences, runs the clustering synchronously and keeps host copies
of the results, as the reference's constructor does (:81-93).  Errors raise
RuntimeError, like the host layer of the reference.
"""
from __future__ import annotations

from typing import Iterable, Sequence

import numpy as np

from . import _lib
from .engine import default_context


def _as_bytes(s) -> bytes:
    if isinstance(s, bytes):
        return s
    if isinstance(s, bytearray):
        return bytes(s)
    return str(s).encode("utf-8")


def pack_strings(strings: Sequence) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Concatenated bytes, uint64 offsets, uint32 lengths (StringSet layout,
    cuda/StringOperations.hcu:77-116)."""
    bs = [_as_bytes(s) for s in strings]
    lens = np.fromiter((len(b) for b in bs), dtype=np.uint32, count=len(bs))
    off = np.zeros(len(bs), np.uint64)
    if len(bs):
        off[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    data = np.frombuffer(b"".join(bs), dtype=np.uint8).copy() if bs else np.zeros(0, np.uint8)
    return data, off, lens


def gene_bitsets(calls: Sequence, genes: Sequence) -> tuple[np.ndarray, int]:
    """R2 on the host through idb_gene_bitsets; returns (words [n, width], width)."""
    width = (len(genes) + 31) // 32
    n = len(calls)
    out = np.zeros((n, max(width, 1)), np.int32)
    cd, co, cl = pack_strings(calls)
    gd, go, gl = pack_strings(genes)
    cd = cd if len(cd) else np.zeros(1, np.uint8)
    gd = gd if len(gd) else np.zeros(1, np.uint8)
    rc = _lib.lib.idb_gene_bitsets(n, cd.ctypes.data, co.ctypes.data, cl.ctypes.data, len(genes), gd.ctypes.data,
                                   go.ctypes.data, gl.ctypes.data, out.ctypes.data, max(width, 1))
    if rc != _lib.IDB_OK:
        raise RuntimeError(f"idb_gene_bitsets failed with code {rc}")
    return out[:, :width], width


def gene_bitsets_gpu(calls: Sequence, genes: Sequence, ctx: "_lib.Context | None" = None,
                     width: int | None = None) -> tuple[np.ndarray, int]:
    """R2 on the device through idb_gene_bitsets_gpu (host arrays in and out);
    returns (words [n, width], width) like gene_bitsets."""
    from .engine import default_context

    ctx = ctx or default_context()
    width = (len(genes) + 31) // 32 if width is None else width
    n = len(calls)
    out = np.zeros((n, max(width, 1)), np.int32)
    cd, co, cl = pack_strings(calls)
    gd, go, gl = pack_strings(genes)
    cd = cd if len(cd) else np.zeros(1, np.uint8)
    gd = gd if len(gd) else np.zeros(1, np.uint8)
    ctx.check(_lib.lib.idb_gene_bitsets_gpu(ctx.handle, _lib.IDB_MEM_HOST, n, cd.ctypes.data, co.ctypes.data,
                                            cl.ctypes.data, len(genes), gd.ctypes.data, go.ctypes.data,
                                            gl.ctypes.data, out.ctypes.data, max(width, 1)))
    return out[:, :width], width


class InterlinkClass:
    """InterlinkClass(InputStrings, n_entries, V_genes_reference, V_gene_variant_count,
    J_genes_reference, J_gene_variant_count, threshold, minPts, use_only_hamming,
    eval_trees)

    ``InputStrings`` is an iterable of (V call, J call, sequence) triples, the
    zip of three string columns that the reference takes as a zip iterator.
    ``abs_k`` (keyword only, not in the reference) selects a fixed edit
    threshold instead of the relative one.
    """

    def __init__(self, InputStrings: Iterable, n_entries: int, V_genes_reference: Sequence,
                 V_gene_variant_count: int, J_genes_reference: Sequence, J_gene_variant_count: int,
                 threshold: float, minPts: int, use_only_hamming: bool, eval_trees: bool, *,
                 abs_k: int = -1, ctx: _lib.Context | None = None):
        rows = list(InputStrings)[:n_entries]
        if len(rows) != n_entries:
            raise RuntimeError(f"InputStrings holds {len(rows)} rows, n_entries is {n_entries}")
        vcalls = [r[0] for r in rows]
        jcalls = [r[1] for r in rows]
        seqs = [r[2] for r in rows]
        vgenes = list(V_genes_reference)[:V_gene_variant_count]
        jgenes = list(J_genes_reference)[:J_gene_variant_count]
        vb, self._v_width = gene_bitsets(vcalls, vgenes)
        jb, jw = gene_bitsets(jcalls, jgenes)
        self._vbitset = np.ascontiguousarray(vb).reshape(-1)
        data, off, lens = pack_strings(seqs)
        n = n_entries
        self._clusters = np.zeros(n, np.uint32)
        self._cores = np.zeros(n, np.int32)
        self._uniques = np.zeros(n, np.uint32)
        self._copies = np.zeros(n, np.uint32)
        self.stats = _lib.IdbStats()
        ctx = ctx or default_context()
        ds = _lib.make_dataset(n, data if len(data) else np.zeros(1, np.uint8), off, lens,
                               np.ascontiguousarray(vb), self._v_width, np.ascontiguousarray(jb), jw,
                               seq_bytes=len(data))
        outs = _lib.make_outputs(self._clusters, self._cores, self._copies, self._uniques)
        prm = _lib.make_params(threshold, abs_k, minPts, use_only_hamming, eval_trees)
        rc = _lib.lib.idb_dbscan(ctx.handle, ds, prm, outs, self.stats)
        if rc != _lib.IDB_OK:
            raise RuntimeError(ctx.last_error())
        self._n_uniques = int(self.stats.n_uniques)
        # the reference's run log (BenchmarkSingleton rows; run_log.write(path) gives its TSV)
        from .runlog import RunLog, log_dbscan

        self.run_log = log_dbscan(RunLog(), n, threshold, minPts, use_only_hamming, self.stats.as_dict(), eval_trees)

    # getters of cuda/Interlink.hpp:59-67
    def get_clusters_iterator(self) -> np.ndarray:
        return self._clusters

    def get_cores_iterator(self) -> np.ndarray:
        return self._cores

    def get_uniques_iterator(self) -> np.ndarray:
        return self._uniques

    def get_copies_iterator(self) -> np.ndarray:
        return self._copies

    def get_n_uniques(self) -> int:
        return self._n_uniques

    def v_gene_container_begin(self) -> np.ndarray:
        return self._vbitset

    def v_gene_container_end(self) -> int:
        return len(self._vbitset)

    def get_v_bitset_width(self) -> int:
        return self._v_width
