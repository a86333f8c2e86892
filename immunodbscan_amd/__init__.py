"""MI355X (gfx950) implementation of immunodbscan's DBSCAN clone-collapsing path.

The compute path is libidb_hip.so (HIP kernels behind the C ABI of
include/idb.h).  Python modules here are thin bindings:

    _lib       ctypes declarations of the C ABI
    interlink  InterlinkClass, the reference's host->device boundary
    engine     run helpers (host arrays, batches, HBM-resident data)
    synth      seeded synthetic repertoires
    configs    the benchmark workloads
"""
__all__ = ["build", "configs", "engine", "interlink", "synth"]
