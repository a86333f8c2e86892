"""R2 gene bitsets on the device (idb_gene_bitsets_gpu) against the oracle's
substring search and the host builder: the reference's gene lists with
IMGT-style calls, short names over a two-letter alphabet (dense overlaps),
calls longer than one LDS window with names across window boundaries, names
that never match (empty, longer than 1024 bytes), device-resident arrays,
padding words, and argument errors."""
from __future__ import annotations

import os
import random

import numpy as np
import pytest

from immunodbscan_amd import _lib
from immunodbscan_amd.interlink import gene_bitsets, gene_bitsets_gpu, pack_strings
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def gene_list(name):
    return [ln for ln in open(os.path.join(GOLDEN, name), "rb").read().split(b"\n") if ln]


def imgt_calls(genes, n, rng):
    calls = []
    for _ in range(n):
        k = rng.randrange(0, 4)
        parts = [b"Homsap " + rng.choice(genes)[:-1] + b"*0%d F" % rng.randrange(1, 9) for _ in range(k)]
        calls.append(b", or ".join(parts))
    return calls


@pytest.mark.gpu
@pytest.mark.parametrize("lst", ["VHgenes.txt", "JHgenes.txt", "DHgenes.txt"])
def test_reference_gene_lists(gpu_ctx, lst):
    rng = random.Random(5)
    genes = gene_list(lst)
    calls = imgt_calls(genes, 2000, rng) + [b"", b"IGH", b"x" * 3000]
    got, width = gene_bitsets_gpu(calls, genes)
    assert width == (len(genes) + 31) // 32
    assert np.array_equal(got, O.gene_bitsets(calls, genes))
    host, _ = gene_bitsets(calls, genes)
    assert np.array_equal(got, host)


@pytest.mark.gpu
def test_dense_overlaps_and_names_that_never_match(gpu_ctx):
    rng = random.Random(7)
    genes = [bytes(rng.choice(b"AB") for _ in range(rng.randrange(1, 10))) for _ in range(70)]
    genes += [b"", b"A" * 1025, b"ABABABAB"]
    calls = [bytes(rng.choice(b"AB") for _ in range(rng.randrange(0, 80))) for _ in range(3000)]
    got, _ = gene_bitsets_gpu(calls, genes)
    assert np.array_equal(got, O.gene_bitsets(calls, genes))


@pytest.mark.gpu
def test_long_calls_across_windows(gpu_ctx):
    rng = random.Random(9)
    genes = [b"IGHV3-23*", b"Q" * 1024, b"R" * 1000, b"WXYZ", b"Z" * 3, b"ST" * 300]
    calls = []
    for _ in range(60):
        L = rng.randrange(900, 5200)
        s = bytearray(rng.choice(b"abcdefgh") for _ in range(L))
        for g in genes:
            if rng.random() < 0.6 and len(g) <= L:
                p = rng.choice([0, 1020, 1023, 1024, 1025, 2040, 2047, 2048, 3070, L - len(g), rng.randrange(0, L - len(g) + 1)])
                p = max(0, min(p, L - len(g)))
                s[p:p + len(g)] = g
        calls.append(bytes(s))
    calls.append(b"Q" * 1023)  # one byte short of the 1024-byte name
    calls.append(b"Q" * 1024)
    got, _ = gene_bitsets_gpu(calls, genes)
    assert np.array_equal(got, O.gene_bitsets(calls, genes))
    assert (got[-2, 0] & 2) == 0 and (got[-1, 0] & 2) == 2


@pytest.mark.gpu
def test_device_arrays_and_padding_words(gpu_ctx):
    from immunodbscan_amd.hip import DeviceBuffer, synchronize

    rng = random.Random(13)
    genes = gene_list("VHgenes.txt")
    calls = imgt_calls(genes, 5000, rng)
    width = 6  # two words past the 104 names
    cd, co, cl = pack_strings(calls)
    gd, go, gl = pack_strings(genes)
    dc, do, dl = DeviceBuffer.from_array(cd), DeviceBuffer.from_array(co), DeviceBuffer.from_array(cl)
    dout = DeviceBuffer.from_array(np.full(len(calls) * width, -1, np.int32))
    ctx = _lib.Context(0)
    ctx.check(_lib.lib.idb_gene_bitsets_gpu(ctx.handle, _lib.IDB_MEM_DEVICE, len(calls), dc.data_ptr(), do.data_ptr(),
                                            dl.data_ptr(), len(genes), gd.ctypes.data, go.ctypes.data, gl.ctypes.data,
                                            dout.data_ptr(), width))
    synchronize()
    got = dout.to_numpy(np.int32, len(calls) * width).reshape(len(calls), width)
    exp = O.gene_bitsets(calls, genes)
    assert np.array_equal(got[:, :exp.shape[1]], exp)
    assert not got[:, exp.shape[1]:].any()
    ctx.close()


@pytest.mark.gpu
def test_empty_inputs_and_argument_errors(gpu_ctx):
    got, width = gene_bitsets_gpu([], [b"IGHV1"])
    assert got.shape == (0, 1)
    got, width = gene_bitsets_gpu([b"IGHV1"], [])
    assert width == 0 and got.shape == (1, 0)
    with pytest.raises(RuntimeError):
        gene_bitsets_gpu([b"IGHV1"], [b"I"] * 40, width=1)
    with pytest.raises(RuntimeError):
        gene_bitsets_gpu([b"IGHV1"], [b"I"], width=257)
