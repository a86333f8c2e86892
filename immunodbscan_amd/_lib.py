"""ctypes binding of libidb_hip.so (declarations in include/idb.h).

The library is the only compute path: importing this module fails loudly when
the built library is missing, and every call raises when the HIP runtime
reports an error.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libidb_hip.so")

IDB_OK = 0
IDB_ERR_ARG = -1
IDB_ERR_HIP = -2
IDB_ERR_OOM = -3
IDB_ERR_UNSUPPORTED = -4
IDB_MEM_HOST = 0
IDB_MEM_DEVICE = 1
IDB_N_STAGES = 8
IDB_FLAG_AUTOTUNE = 1
IDB_FLAG_GRAPH = 2
IDB_FLAG_FORCE_LARGE_PATHS = 4
IDB_FLAG_SPLIT_TILES = 8
IDB_NOALLOC_MAX_ROWS = 65536

STAGE_NAMES = [
    "Identify Duplicates",
    "Reduction of duplicate information tree structure",
    "Determine copies",
    "Identify core points",
    "Grow clusters",
    "Reduction of cluster information tree structure",
    "Re-number clusters and identify noise",
    "Clustering core process",
]

EXPORTED = [
    "idb_create",
    "idb_destroy",
    "idb_last_error",
    "idb_gene_bitsets",
    "idb_gene_bitsets_gpu",
    "idb_dbscan",
    "idb_dbscan_noalloc",
    "idb_dbscan_batch",
    "idb_pair_distances",
    "idb_neighbour_pairs",
    "idb_encoder_create",
    "idb_encoder_create_positions",
    "idb_encoder_positions",
    "idb_encoder_destroy",
    "idb_encoder_last_error",
    "idb_encoder_param_count",
    "idb_encoder_get_params",
    "idb_encoder_set_params",
    "idb_encoder_embed",
    "idb_encoder_loss_grad",
    "idb_encoder_train",
    "idb_dbscan_learned",
]

IDB_ENC_PARAMS = 193984


class IdbDataset(C.Structure):
    _fields_ = [
        ("n_rows", C.c_uint32),
        ("mem", C.c_uint32),
        ("seq", C.c_void_p),
        ("seq_bytes", C.c_uint64),
        ("off", C.c_void_p),
        ("len", C.c_void_p),
        ("vbits", C.c_void_p),
        ("v_width", C.c_uint32),
        ("jbits", C.c_void_p),
        ("j_width", C.c_uint32),
    ]


class IdbParams(C.Structure):
    _fields_ = [
        ("eps", C.c_float),
        ("abs_k", C.c_int32),
        ("min_pts", C.c_uint32),
        ("use_hamming", C.c_uint32),
        ("tree_stats", C.c_uint32),
        ("debug_hash_bits", C.c_uint32),
        ("edge_capacity", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class IdbOutputs(C.Structure):
    _fields_ = [
        ("mem", C.c_uint32),
        ("clone_id", C.c_void_p),
        ("core", C.c_void_p),
        ("copies", C.c_void_p),
        ("uniques", C.c_void_p),
        ("neighbour_count", C.c_void_p),
    ]


class IdbStats(C.Structure):
    _fields_ = [
        ("n_rows", C.c_uint32),
        ("n_uniques", C.c_uint32),
        ("n_cores", C.c_uint32),
        ("n_clusters", C.c_uint32),
        ("n_noise", C.c_uint32),
        ("max_len", C.c_uint32),
        ("alphabet_size", C.c_uint32),
        ("n_buckets", C.c_uint32),
        ("fuse_from_edges", C.c_uint32),
        ("band_filter", C.c_uint32),
        ("bucket_entries", C.c_uint64),
        ("tiles", C.c_uint64),
        ("gated_pairs", C.c_uint64),
        ("neighbour_pairs", C.c_uint64),
        ("fuse_pairs", C.c_uint64),
        ("gated_letters", C.c_uint64),
        ("myers_pairs", C.c_uint64),
        ("myers_letters", C.c_uint64),
        ("count_kernel_ms", C.c_float),
        ("fuse_kernel_ms", C.c_float),
        ("tree_depth", C.c_uint32 * 51),
        ("stage_ms", C.c_float * IDB_N_STAGES),
        ("tuned_wpb", C.c_uint32),
        ("tuned_grid_mul", C.c_uint32),
        ("tuned_xcd_order", C.c_uint32),
        ("tuned_trials", C.c_uint32),
        ("graph_used", C.c_uint32),
        ("encoder_ms", C.c_float),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("tree_depth", "stage_ms")}
        d["tree_depth"] = list(self.tree_depth)
        d["stage_ms"] = {n: float(v) for n, v in zip(STAGE_NAMES, self.stage_ms)}
        return d


class IdbTrainParams(C.Structure):
    _fields_ = [
        ("epochs", C.c_uint32),
        ("batch_pairs", C.c_uint32),
        ("lr", C.c_float),
        ("margin", C.c_float),
        ("seed", C.c_uint64),
        ("log_every", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class IdbTrainStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint32),
        ("first_loss", C.c_float),
        ("last_loss", C.c_float),
        ("mean_sampled_loss_last_epoch", C.c_float),
        ("train_ms", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m immunodbscan_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    lib.idb_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.idb_create.restype = C.c_int
    lib.idb_destroy.argtypes = [vp]
    lib.idb_destroy.restype = None
    lib.idb_last_error.argtypes = [vp]
    lib.idb_last_error.restype = C.c_char_p
    lib.idb_gene_bitsets.argtypes = [u32, vp, vp, vp, u32, vp, vp, vp, vp, u32]
    lib.idb_gene_bitsets.restype = C.c_int
    lib.idb_gene_bitsets_gpu.argtypes = [vp, u32, u32, vp, vp, vp, u32, vp, vp, vp, vp, u32]
    lib.idb_gene_bitsets_gpu.restype = C.c_int
    lib.idb_dbscan.argtypes = [vp, C.POINTER(IdbDataset), C.POINTER(IdbParams), C.POINTER(IdbOutputs),
                               C.POINTER(IdbStats)]
    lib.idb_dbscan.restype = C.c_int
    lib.idb_dbscan_noalloc.argtypes = [vp, C.POINTER(IdbDataset), C.POINTER(IdbParams), C.POINTER(IdbOutputs),
                                       C.POINTER(IdbStats)]
    lib.idb_dbscan_noalloc.restype = C.c_int
    lib.idb_dbscan_batch.argtypes = [vp, C.POINTER(IdbDataset), vp, u32, C.POINTER(IdbParams),
                                     C.POINTER(IdbOutputs), vp, C.POINTER(IdbStats)]
    lib.idb_dbscan_batch.restype = C.c_int
    DS, PR, OUT, ST = C.POINTER(IdbDataset), C.POINTER(IdbParams), C.POINTER(IdbOutputs), C.POINTER(IdbStats)
    f32 = C.c_float
    sigs = {
        "idb_pair_distances": ([vp, DS, vp, u64, u32, vp], C.c_int),
        "idb_neighbour_pairs": ([vp, DS, PR, vp, u64, C.POINTER(u64)], C.c_int),
        "idb_encoder_create": ([C.c_int, vp, u64, C.POINTER(vp)], C.c_int),
        "idb_encoder_create_positions": ([C.c_int, vp, u32, u64, C.POINTER(vp)], C.c_int),
        "idb_encoder_positions": ([vp], u32),
        "idb_encoder_destroy": ([vp], None),
        "idb_encoder_last_error": ([vp], C.c_char_p),
        "idb_encoder_param_count": ([], u32),
        "idb_encoder_get_params": ([vp, vp], C.c_int),
        "idb_encoder_set_params": ([vp, vp], C.c_int),
        "idb_encoder_embed": ([vp, DS, vp], C.c_int),
        "idb_encoder_loss_grad": ([vp, DS, vp, vp, u32, f32, C.POINTER(f32), vp], C.c_int),
        "idb_encoder_train": ([vp, DS, vp, vp, u64, C.POINTER(IdbTrainParams), C.POINTER(IdbTrainStats)], C.c_int),
        "idb_dbscan_learned": ([vp, vp, f32, DS, PR, OUT, ST], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def _ptr(a) -> int | None:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor on the device


class IdbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"idb error {code}: {msg}")
        self.code = code


class Context:
    """An idb_ctx on one HIP device (stream + reusable device arena)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.idb_create(C.byref(h), device)
        if rc != IDB_OK:
            raise IdbError(rc, f"idb_create failed on device {device} (no usable HIP device?)")
        self._h = h
        self.device = device

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.idb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def last_error(self) -> str:
        return lib.idb_last_error(self._h).decode(errors="replace")

    def check(self, rc: int) -> None:
        if rc != IDB_OK:
            raise IdbError(rc, self.last_error())


def make_params(eps: float = 0.15, abs_k: int = -1, min_pts: int = 2, use_hamming: bool = False,
                tree_stats: bool = False, debug_hash_bits: int = 0, edge_capacity: int = 0,
                flags: int = 0) -> IdbParams:
    p = IdbParams()
    p.eps = float(np.float32(eps))
    p.abs_k = int(abs_k)
    p.min_pts = int(min_pts) & 0xFFFFFFFF
    p.use_hamming = 1 if use_hamming else 0
    p.tree_stats = 1 if tree_stats else 0
    p.debug_hash_bits = int(debug_hash_bits)
    p.edge_capacity = int(edge_capacity)
    p.flags = int(flags)
    return p


def make_dataset(n_rows: int, seq, off, length, vbits, v_width: int, jbits, j_width: int,
                 mem: int = IDB_MEM_HOST, seq_bytes: int | None = None) -> IdbDataset:
    d = IdbDataset()
    d.n_rows = int(n_rows)
    d.mem = mem
    d.seq = _ptr(seq)
    if seq_bytes is None:
        seq_bytes = int(seq.nbytes if isinstance(seq, np.ndarray) else seq.numel() * seq.element_size())
    d.seq_bytes = int(seq_bytes)
    d.off = _ptr(off)
    d.len = _ptr(length)
    d.vbits = _ptr(vbits)
    d.v_width = int(v_width)
    d.jbits = _ptr(jbits)
    d.j_width = int(j_width)
    return d


def make_outputs(clone_id, core, copies, uniques, neighbour_count=None, mem: int = IDB_MEM_HOST) -> IdbOutputs:
    o = IdbOutputs()
    o.mem = mem
    o.clone_id = _ptr(clone_id)
    o.core = _ptr(core)
    o.copies = _ptr(copies)
    o.uniques = _ptr(uniques)
    o.neighbour_count = _ptr(neighbour_count)
    return o
