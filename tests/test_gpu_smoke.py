"""First GPU parity checks: the HIP path against the oracle on small inputs."""
import numpy as np
import pytest

from immunodbscan_amd import configs, engine, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cmp(rep, cl, fast=True):
    got = engine.dbscan(rep, cl)
    exp = O.dbscan(O.Data.of(rep), O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming), fast=fast)
    for k in ("copies", "core", "uniques", "clone_id"):
        assert np.array_equal(got.__dict__[k] if k != "clone_id" else got.clone_id, exp[k]), k
    assert np.array_equal(got.counts, exp["counts"])
    assert got.n_uniques == exp["n_uniques"]
    return got, exp


def test_c1_hamming(gpu_ctx):
    rep = synth.generate(configs.c1_params())
    print("c1", rep.n, flush=True)
    _cmp(rep, configs.C1, fast=False)


def test_c2_small_levenshtein(gpu_ctx):
    rep = synth.generate(configs.c2_params(n_rows=5000, n_seeds=500))
    print("c2 small", rep.n, flush=True)
    _cmp(rep, configs.C2, fast=False)


def test_c1_levenshtein_relative(gpu_ctx):
    rep = synth.generate(configs.c1_params(n_rows=3000, n_seeds=300, seed=11))
    cl = configs.Clustering(eps=0.1, abs_k=-1, min_pts=3, use_hamming=False)
    print("c1 lev", rep.n, flush=True)
    _cmp(rep, cl, fast=False)
