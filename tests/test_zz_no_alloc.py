"""The allocation-free entry point idb_dbscan_noalloc, through the C ABI.

Its results must equal idb_dbscan's bit for bit (clone ids, core flags,
copies, uniques, exact counts, statistics) on two calls in a row, the
second with a smaller input (the scratch array keeps the first call's
contents), with host- and device-resident arrays; inputs past its limits fail
with a clear error and leave the context usable.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from immunodbscan_amd import _lib, configs, engine, synth

pytestmark = pytest.mark.gpu
FIELDS = ("clone_id", "core", "copies", "uniques", "counts")


def noalloc(rep, cl, ctx):
    """idb_dbscan_noalloc with host arrays, as engine.dbscan calls idb_dbscan."""
    n = rep.n
    out = {"clone_id": np.zeros(n, np.uint32), "core": np.zeros(n, np.int32), "copies": np.zeros(n, np.uint32),
           "uniques": np.zeros(n, np.uint32), "counts": np.zeros(n, np.uint32)}
    ds = engine._dataset(rep)
    outs = _lib.make_outputs(out["clone_id"], out["core"], out["copies"], out["uniques"], out["counts"])
    st = _lib.IdbStats()
    prm = engine._params(cl, False, 0)
    rc = _lib.lib.idb_dbscan_noalloc(ctx.handle, ds, prm, outs, st)
    return rc, out, st.as_dict()


def same(a, b, st_a, st_b, tag):
    for k in FIELDS:
        assert np.array_equal(a[k], getattr(b, k)), (tag, k)
    for k in ("n_uniques", "n_cores", "n_clusters", "n_noise", "gated_pairs", "neighbour_pairs", "tiles",
              "bucket_entries"):
        assert st_a[k] == st_b[k], (tag, k, st_a[k], st_b[k])


def test_noalloc_equals_dbscan_two_calls_second_smaller(gpu_ctx):
    ctx = _lib.Context(0)
    cases = [
        ("c2 60k", synth.generate(configs.c2_params(n_rows=60_000, n_seeds=6_000, seed=71)), configs.C2),
        ("c2 12k", synth.generate(configs.c2_params(n_rows=12_000, n_seeds=1_200, seed=72)), configs.C2),
        ("c1 2k", synth.generate(configs.c1_params()), configs.C1),
        ("c5 8k", synth.generate(configs.c5_params(n_rows=8_000, n_seeds=80, seed=73)), configs.C5),
    ]
    for tag, rep, cl in cases:
        print(f"case noalloc {tag}: n={rep.n}", flush=True)
        ref = engine.dbscan(rep, cl, ctx=ctx)
        rc, got, st = noalloc(rep, cl, ctx)
        assert rc == _lib.IDB_OK, (tag, rc, ctx.last_error())
        same(got, ref, st, ref.stats, tag)
    ctx.close()


def test_noalloc_device_resident_arrays(gpu_ctx):
    ctx = _lib.Context(0)
    for n, seed in ((40_000, 74), (5_000, 75)):
        rep = synth.generate(configs.c2_params(n_rows=n, n_seeds=n // 10, seed=seed))
        print(f"case noalloc device n={n}", flush=True)
        run = engine.DeviceRun(rep, configs.C2, ctx=ctx)
        rc = _lib.lib.idb_dbscan_noalloc(ctx.handle, run.ds, run.prm, run.outs, run.stats)
        assert rc == _lib.IDB_OK, ctx.last_error()
        got = run.result()
        ref = engine.dbscan(rep, configs.C2, ctx=ctx)
        for k in FIELDS:
            assert np.array_equal(getattr(got, k), getattr(ref, k)), (n, k)
    ctx.close()


def test_noalloc_limits_are_clean_errors(gpu_ctx):
    ctx = _lib.Context(0)
    big = synth.generate(configs.c2_params(n_rows=_lib.IDB_NOALLOC_MAX_ROWS + 1, n_seeds=6_000, seed=76))
    rc, _, _ = noalloc(big, configs.C2, ctx)
    assert rc == _lib.IDB_ERR_UNSUPPORTED and "at most" in ctx.last_error()
    rep = synth.generate(configs.c2_params(n_rows=3_000, n_seeds=300, seed=77))
    prm = engine._params(configs.C2, False, 0, 0, _lib.IDB_FLAG_GRAPH)
    n = rep.n
    bufs = [np.zeros(n, np.uint32) for _ in range(4)]
    outs = _lib.make_outputs(*bufs)
    st = _lib.IdbStats()
    assert _lib.lib.idb_dbscan_noalloc(ctx.handle, engine._dataset(rep), prm, outs, st) == _lib.IDB_ERR_ARG
    # the context still works, on both entry points
    rc, got, st2 = noalloc(rep, configs.C2, ctx)
    assert rc == _lib.IDB_OK
    ref = engine.dbscan(rep, configs.C2, ctx=ctx)
    same(got, ref, st2, ref.stats, "after errors")
    ctx.close()
