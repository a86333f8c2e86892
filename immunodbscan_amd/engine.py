"""Run helpers over the C ABI: one repertoire, a batch, or device-resident data."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .configs import Clustering
from .synth import Repertoire


@dataclass
class Result:
    clone_id: np.ndarray
    core: np.ndarray
    copies: np.ndarray
    uniques: np.ndarray
    n_uniques: int | np.ndarray
    counts: np.ndarray | None
    stats: dict = field(default_factory=dict)


_default_ctx: _lib.Context | None = None


def default_context() -> _lib.Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = _lib.Context(0)
    return _default_ctx


def _params(cl: Clustering, tree_stats: bool, debug_hash_bits: int, edge_capacity: int = 0,
            flags: int = 0) -> _lib.IdbParams:
    return _lib.make_params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming, tree_stats, debug_hash_bits,
                            edge_capacity, flags)


def _dataset(rep: Repertoire) -> _lib.IdbDataset:
    seq = rep.seq if len(rep.seq) else np.zeros(1, np.uint8)
    return _lib.make_dataset(rep.n, seq, rep.off, rep.length, np.ascontiguousarray(rep.vbits), rep.vw,
                             np.ascontiguousarray(rep.jbits), rep.jw, seq_bytes=len(rep.seq))


def dbscan(rep: Repertoire, cl: Clustering, ctx: _lib.Context | None = None, counts: bool = True,
           tree_stats: bool = False, debug_hash_bits: int = 0, edge_capacity: int = 0, flags: int = 0) -> Result:
    ctx = ctx or default_context()
    n = rep.n
    cid = np.zeros(n, np.uint32)
    core = np.zeros(n, np.int32)
    cop = np.zeros(n, np.uint32)
    unq = np.zeros(n, np.uint32)
    cnt = np.zeros(n, np.uint32) if counts else None
    ds = _dataset(rep)
    outs = _lib.make_outputs(cid, core, cop, unq, cnt)
    st = _lib.IdbStats()
    prm = _params(cl, tree_stats, debug_hash_bits, edge_capacity, flags)
    ctx.check(_lib.lib.idb_dbscan(ctx.handle, ds, prm, outs, st))
    return Result(cid, core, cop, unq, int(st.n_uniques), cnt, st.as_dict())


def dbscan_batch(rep: Repertoire, cl: Clustering, ctx: _lib.Context | None = None, counts: bool = True,
                 edge_capacity: int = 0, flags: int = 0) -> Result:
    ctx = ctx or default_context()
    n = rep.n
    starts = np.ascontiguousarray(rep.rep_start, dtype=np.uint32)
    nr = len(starts) - 1
    cid = np.zeros(n, np.uint32)
    core = np.zeros(n, np.int32)
    cop = np.zeros(n, np.uint32)
    unq = np.zeros(n, np.uint32)
    cnt = np.zeros(n, np.uint32) if counts else None
    nu = np.zeros(max(nr, 1), np.uint32)
    ds = _dataset(rep)
    outs = _lib.make_outputs(cid, core, cop, unq, cnt)
    st = _lib.IdbStats()
    prm = _params(cl, False, 0, edge_capacity, flags)
    ctx.check(_lib.lib.idb_dbscan_batch(ctx.handle, ds, starts.ctypes.data, nr, prm, outs, nu.ctypes.data, st))
    return Result(cid, core, cop, unq, nu[:nr], cnt, st.as_dict())


class HostRun:
    """A repertoire in host memory, clustered the way the reference's
    boundary does it (InterlinkClass, cuda/Interlink.cu:34-95): each call
    copies the rows to the device, clusters them and copies clone ids, core
    flags, copies and uniques back into host arrays.  Buffers are kept across
    calls."""

    def __init__(self, rep: Repertoire, cl: Clustering, ctx: _lib.Context | None = None, counts: bool = False,
                 flags: int = 0):
        self.ctx = ctx or default_context()
        self.rep = rep
        n = rep.n
        self.cid = np.zeros(n, np.uint32)
        self.core = np.zeros(n, np.int32)
        self.cop = np.zeros(n, np.uint32)
        self.unq = np.zeros(n, np.uint32)
        self.cnt = np.zeros(n, np.uint32) if counts else None
        # the arrays the dataset points at are kept alive with the run
        self._keep = (rep.seq if len(rep.seq) else np.zeros(1, np.uint8), np.ascontiguousarray(rep.off),
                      np.ascontiguousarray(rep.length), np.ascontiguousarray(rep.vbits),
                      np.ascontiguousarray(rep.jbits))
        seq, off, ln, vb, jb = self._keep
        self.ds = _lib.make_dataset(n, seq, off, ln, vb, rep.vw, jb, rep.jw, seq_bytes=len(rep.seq))
        self.outs = _lib.make_outputs(self.cid, self.core, self.cop, self.unq, self.cnt)
        self.prm = _params(cl, False, 0, 0, flags)
        self.stats = _lib.IdbStats()

    def __call__(self) -> _lib.IdbStats:
        self.ctx.check(_lib.lib.idb_dbscan(self.ctx.handle, self.ds, self.prm, self.outs, self.stats))
        return self.stats

    def result(self) -> Result:
        return Result(self.cid, self.core, self.cop, self.unq, int(self.stats.n_uniques), self.cnt,
                      self.stats.as_dict())


class DeviceRun:
    """A repertoire resident in HBM with device-side outputs, for timing the
    path without host copies.  Buffers come from immunodbscan_amd.hip."""

    def __init__(self, rep: Repertoire, cl: Clustering, ctx: _lib.Context | None = None, batch: bool = False,
                 flags: int = 0, encoder=None, min_sim: float = 0.0):
        from .hip import DeviceBuffer, synchronize

        self.ctx = ctx or default_context()
        dev = self.ctx.device
        up = lambda a: DeviceBuffer.from_array(a, dev)  # noqa: E731
        self.rep = rep
        self.seq = up(rep.seq if len(rep.seq) else np.zeros(1, np.uint8))
        self.off = up(rep.off)
        self.len = up(rep.length)
        self.vb = up(rep.vbits)
        self.jb = up(rep.jbits)
        n = rep.n
        z32 = lambda: DeviceBuffer.zeros(4 * n, dev)  # noqa: E731
        self.cid, self.core, self.cop, self.unq, self.cnt = z32(), z32(), z32(), z32(), z32()
        self.ds = _lib.make_dataset(n, self.seq, self.off, self.len, self.vb, rep.vw, self.jb, rep.jw,
                                    mem=_lib.IDB_MEM_DEVICE, seq_bytes=len(rep.seq))
        self.outs = _lib.make_outputs(self.cid, self.core, self.cop, self.unq, self.cnt, mem=_lib.IDB_MEM_DEVICE)
        self.prm = _params(cl, False, 0, 0, flags)
        self.batch = batch
        self.starts = np.ascontiguousarray(rep.rep_start, dtype=np.uint32)
        self.nu = np.zeros(max(len(self.starts) - 1, 1), np.uint32)
        self.stats = _lib.IdbStats()
        self.encoder = encoder
        self.min_sim = float(min_sim)
        synchronize()

    def __call__(self) -> _lib.IdbStats:
        if self.encoder is not None:
            rc = _lib.lib.idb_dbscan_learned(self.ctx.handle, self.encoder.handle, self.min_sim, self.ds, self.prm,
                                             self.outs, self.stats)
        elif self.batch:
            rc = _lib.lib.idb_dbscan_batch(self.ctx.handle, self.ds, self.starts.ctypes.data, len(self.starts) - 1,
                                           self.prm, self.outs, self.nu.ctypes.data, self.stats)
        else:
            rc = _lib.lib.idb_dbscan(self.ctx.handle, self.ds, self.prm, self.outs, self.stats)
        self.ctx.check(rc)
        return self.stats

    def result(self) -> Result:
        n = self.rep.n
        return Result(self.cid.to_numpy(np.uint32, n), self.core.to_numpy(np.int32, n),
                      self.cop.to_numpy(np.uint32, n), self.unq.to_numpy(np.uint32, n),
                      int(self.stats.n_uniques) if not self.batch else self.nu.copy(),
                      self.cnt.to_numpy(np.uint32, n), self.stats.as_dict())
