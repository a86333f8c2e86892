"""CPU-only checks of the native libraries: the C ABI exports and struct
layouts, the host gene-bitset builder, the generator's determinism, and the
Myers / packed-Hamming primitives compiled for the host."""
from __future__ import annotations

import ctypes
import json
import os
import random
import re
import subprocess
import tempfile

import numpy as np
import pytest

from immunodbscan_amd import configs, synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "idb.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|const char \*)\s*(idb_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from immunodbscan_amd import _lib

    names = declared_functions()
    assert names == sorted(_lib.EXPORTED)
    so = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(so, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (idb_\w+)", out))
    assert set(names) <= exported


def test_struct_layouts_match_header():
    from immunodbscan_amd import _lib

    src = """
#include <stdio.h>
#include <stddef.h>
#include "idb.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(idb_dataset), sizeof(idb_params), sizeof(idb_outputs), sizeof(idb_stats));
  printf("%zu %zu %zu\\n", offsetof(idb_stats, tree_depth), offsetof(idb_stats, stage_ms), offsetof(idb_dataset, j_width));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(v) for v in lines[0].split()]
    assert sizes == [ctypes.sizeof(_lib.IdbDataset), ctypes.sizeof(_lib.IdbParams), ctypes.sizeof(_lib.IdbOutputs),
                     ctypes.sizeof(_lib.IdbStats)]
    offs = [int(v) for v in lines[1].split()]
    assert offs == [_lib.IdbStats.tree_depth.offset, _lib.IdbStats.stage_ms.offset, _lib.IdbDataset.j_width.offset]


def test_host_gene_bitsets_match_oracle():
    from immunodbscan_amd.interlink import gene_bitsets

    rng = random.Random(11)
    vh = [ln for ln in open(os.path.join(GOLDEN, "VHgenes.txt"), "rb").read().split(b"\n") if ln]
    calls = []
    for _ in range(300):
        k = rng.randrange(0, 3)
        parts = [b"Homsap " + rng.choice(vh)[:-1] + b"*0%d F" % rng.randrange(1, 9) for _ in range(k)]
        calls.append(b", or ".join(parts))
    calls += [b"", b"IGHV", b"x" * 3000]
    got, width = gene_bitsets(calls, vh)
    assert width == 4
    assert np.array_equal(got, O.gene_bitsets(calls, vh))
    # random alphabet, short names (many overlaps)
    genes = [bytes(rng.choice(b"AB") for _ in range(rng.randrange(1, 5))) for _ in range(70)]
    calls = [bytes(rng.choice(b"AB") for _ in range(rng.randrange(0, 9))) for _ in range(200)]
    got, width = gene_bitsets(calls, genes + [b"", b"A" * 1025])
    assert np.array_equal(got, O.gene_bitsets(calls, genes + [b"", b"A" * 1025]))


def test_generator_is_deterministic_and_matches_fixture():
    from tests.golden.make_golden import digest, instances

    sums = json.load(open(os.path.join(GOLDEN, "synth_checksums.json")))
    for name, p in instances().items():
        assert digest(synth.generate(p)) == sums[name], name


def test_generator_shapes():
    rep = synth.generate(configs.c2_params(n_rows=20_000, n_seeds=2_000))
    assert rep.n == 20_000 and rep.vw == 2 and rep.jw == 1
    assert rep.length.min() >= 5 and rep.length.max() <= 28
    assert set(np.unique(rep.seq).tobytes()) <= set(b"ACDEFGHIKLMNPQRSTVWY")
    ng = np.array([bin(int(np.uint32(w))).count("1") for w in rep.vbits.view(np.uint32).ravel()]).reshape(rep.n, 2)
    per_row = ng.sum(1)
    assert per_row.min() == 1 and per_row.max() == 2
    assert 0.15 < (per_row == 2).mean() < 0.25
    rep1 = synth.generate(configs.c1_params())
    assert rep1.n == 2000 and set(np.unique(rep1.length)) <= set(range(30, 76, 3))


def test_oracle_c1_fixture_still_holds():
    rep = synth.generate(configs.c1_params())
    cl = configs.C1
    r = O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming), fast=True)
    fx = np.load(os.path.join(GOLDEN, "oracle_c1.npz"))
    for k in fx.files:
        assert np.array_equal(fx[k], r[k]), k


def test_pair_primitives_on_host():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "t")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "test_pairdist.cpp"), "-o", exe],
                       check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "OK mismatches=0" in r.stdout


@pytest.mark.parametrize("flags", [["-O2"], ["-O1", "-g", "-fsanitize=thread"]], ids=["O2", "tsan"])
def test_copy_pool_parts(flags):
    """The host copy pool behind the staged input transfer (one pass, parts handed over in order)."""
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "t")
        b = subprocess.run(["g++", *flags, "-std=c++17", "-pthread", os.path.join(ROOT, "tests", "cpp", "test_copy_pool.cpp"),
                            "-o", exe], capture_output=True, text=True)
        if b.returncode != 0 and "-fsanitize=thread" in flags:
            pytest.skip("no thread sanitizer in this toolchain: " + b.stderr[-200:])
        assert b.returncode == 0, b.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK copy pool" in r.stdout

