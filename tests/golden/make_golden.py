"""Writes the golden fixtures of this directory.

    synth_checksums.json  SHA-256 of every array the generator produces for
                          small instances of the benchmark workloads, so that
                          a machine that regenerates them can tell whether it
                          produced the same rows
    oracle_c1.npz         oracle outputs (clone ids, cores, copies, uniques,
                          exact counts) for the 2,000-row Hamming workload

The gene lists (VHgenes.txt, JHgenes.txt, DHgenes.txt) are the data files the
reference ships under data/; they are fixtures for the R2 bit order.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from immunodbscan_amd import configs, synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(rep) -> dict:
    out = {}
    for k in ("seq", "off", "length", "vbits", "jbits", "clone_of"):
        out[k] = hashlib.sha256(np.ascontiguousarray(getattr(rep, k)).tobytes()).hexdigest()
    out["n"] = rep.n
    return out


def instances():
    return {
        "c1": configs.c1_params(),
        "c2_small": configs.c2_params(n_rows=20_000, n_seeds=2_000),
        "c5_small": configs.c5_params(n_rows=20_000, n_seeds=100),
        "c3_rep0": configs.c1_params(n_rows=int(configs.c3_sizes()[0]), n_seeds=max(1, int(configs.c3_sizes()[0]) // 10),
                                     seed=3 * 1000003, v_zipf=1.1),
    }


def main():
    sums = {name: digest(synth.generate(p)) for name, p in instances().items()}
    with open(os.path.join(HERE, "synth_checksums.json"), "w") as f:
        json.dump(sums, f, indent=1, sort_keys=True)
    rep = synth.generate(configs.c1_params())
    cl = configs.C1
    r = O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming), fast=False)
    np.savez_compressed(os.path.join(HERE, "oracle_c1.npz"),
                        **{k: r[k] for k in ("clone_id", "core", "copies", "uniques", "counts")})
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
