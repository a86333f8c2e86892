"""The five benchmark workloads, as generator settings plus clustering settings.

c1  2,000 nt CDR3s, 200 seed clonotypes (geometric sizes, mean 10), lengths
    30-75 nt in steps of 3, 2% substitutions, 50 V genes (second call p=0.2),
    6 J genes; Hamming, eps = 0.15 relative, minPts = 2.  Seed 1.
c2  200,000 aa CDR3s, 20,000 seeds (lognormal mu=1 sigma=1.2, scaled to the
    row count), lengths round(N(15, 3)) clipped to 8-25, 1-3 edits per member,
    64 V genes Zipf(1.1) (second call p=0.2), 8 J genes; Levenshtein, t = 2
    edits, minPts = 3.  Seed 2.
c3  512 repertoires, sizes lognormal (median 5,000, sigma 1.82, clipped to
    200-200,000: sigma chosen so that the seed-3 draw totals about 10M rows,
    10,018,569), each generated as c1 with size/10 seeds and Zipf(1.1) V
    usage.  Seed 3 (repertoire r uses seed 3 * 1000003 + r).
c4  2,000,000 aa CDR3s generated as c2 with 200,000 seeds (the encoder of the
    learned filter works on amino acids).  Seed 4.
c5  20,000,000 nt CDR3s, 100,000 seeds (lognormal, scaled), lengths 24-150 nt,
    3% substitutions and 0.5% indels per letter, 50 V genes Zipf(1.1), 6 J
    genes; Levenshtein, t = 3 edits, minPts = 3.  Seed 5.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import synth


@dataclass
class Clustering:
    eps: float = 0.15
    abs_k: int = -1
    min_pts: int = 2
    use_hamming: bool = False


def c1_params(n_rows: int = 2000, n_seeds: int = 200, seed: int = 1, v_zipf: float = 0.0) -> synth.SynthParams:
    p = synth.default_params()
    p.seed = seed
    p.n_rows = n_rows
    p.n_seeds = n_seeds
    p.alphabet = synth.NT
    p.len_mode = synth.LEN_UNIFORM
    p.len_lo, p.len_hi, p.len_step = 30, 75, 3
    p.size_mode = synth.SIZE_GEOMETRIC
    p.size_p = 0.1
    p.mut_mode = synth.MUT_SUB
    p.sub_rate = 0.02
    p.n_v, p.n_j = 50, 6
    p.v_zipf_s = v_zipf
    p.v_ambig = 0.2
    p.shuffle = 1
    return p


C1 = Clustering(eps=0.15, abs_k=-1, min_pts=2, use_hamming=True)


def c2_params(n_rows: int = 200_000, n_seeds: int = 20_000, seed: int = 2) -> synth.SynthParams:
    p = synth.default_params()
    p.seed = seed
    p.n_rows = n_rows
    p.n_seeds = n_seeds
    p.alphabet = synth.AA
    p.len_mode = synth.LEN_NORMAL
    p.len_lo, p.len_hi = 8, 25
    p.len_mean, p.len_sd = 15.0, 3.0
    p.size_mode = synth.SIZE_LOGNORMAL
    p.size_mu, p.size_sigma = 1.0, 1.2
    p.mut_mode = synth.MUT_EDITS
    p.k_lo, p.k_hi = 1, 3
    p.n_v, p.n_j = 64, 8
    p.v_zipf_s = 1.1
    p.v_ambig = 0.2
    p.shuffle = 1
    return p


C2 = Clustering(eps=0.0, abs_k=2, min_pts=3, use_hamming=False)


def c3_sizes(n_reps: int = 512, seed: int = 3) -> np.ndarray:
    return synth.lognormal_sizes(seed, n_reps, 5000.0, 1.82, 200, 200_000)


def c3_repertoires(n_reps: int = 512, seed: int = 3, scale: float = 1.0) -> list[synth.Repertoire]:
    sizes = c3_sizes(n_reps, seed)
    reps = []
    for r, s in enumerate(sizes):
        s = max(2, int(round(int(s) * scale)))
        reps.append(synth.generate(c1_params(n_rows=s, n_seeds=max(1, s // 10), seed=seed * 1000003 + r,
                                             v_zipf=1.1)))
    return reps


C3 = C1


def c4_params(n_rows: int = 2_000_000, seed: int = 4) -> synth.SynthParams:
    return c2_params(n_rows=n_rows, n_seeds=max(1, n_rows // 10), seed=seed)


C4 = C2


def c5_params(n_rows: int = 20_000_000, n_seeds: int = 100_000, seed: int = 5) -> synth.SynthParams:
    p = synth.default_params()
    p.seed = seed
    p.n_rows = n_rows
    p.n_seeds = n_seeds
    p.alphabet = synth.NT
    p.len_mode = synth.LEN_UNIFORM
    p.len_lo, p.len_hi, p.len_step = 24, 150, 1
    p.size_mode = synth.SIZE_LOGNORMAL
    p.size_mu, p.size_sigma = 1.0, 1.2
    p.mut_mode = synth.MUT_SUB_INDEL
    p.sub_rate, p.indel_rate = 0.03, 0.005
    p.n_v, p.n_j = 50, 6
    p.v_zipf_s = 1.1
    p.v_ambig = 0.2
    p.shuffle = 1
    return p


C5 = Clustering(eps=0.0, abs_k=3, min_pts=3, use_hamming=False)
