#!/usr/bin/env python3
"""Exact DBSCAN clone collapsing on one MI355X.

Workload (default): one repertoire of 200,000 amino-acid CDR3s from 20,000
seed clonotypes (lognormal clone sizes, 1-3 edits per member, lengths
N(15, 3) clipped to 8-25), 64 V genes (Zipf 1.1, second call p=0.2) and 8 J
genes; neighbours are pairs with a common V and J gene and Levenshtein
distance <= 2; minPts = 3.  A step is one full clustering call as the
reference's boundary makes it (InterlinkClass): rows in host memory at entry,
copied to the device, clustered (duplicates, core points, clusters, labels),
and clone ids, core flags, copies and uniques copied back to host memory.
The same call with rows already in HBM and labels left there is reported as
device_resident.

Metric: pairs/sec under the reference's all-pairs work model, N(N-1)/2 per
step divided by the step time; seqs clustered/sec is reported beside it.
The CPU baseline is the oracle's literal all-pairs restatement of the
reference, run on every core this process may use, on the whole workload.

Beside the headline exact line the JSON carries:
  c1        the 2,000-row Hamming workload, GPU and the literal CPU oracle
            (median of 3) side by side;
  learned   the learned candidate filter on the same workload and on a
            2,000,000-row repertoire of the same generator: the encoder is
            trained on the device (untimed) on pairs labelled by the exact
            kernels of a separate repertoire, its threshold is tuned for
            99.9% neighbour-pair recall on a third one, and the timed step is
            the whole learned clustering call (embedding included); recall,
            ARI and the speedup over the exact path are reported;
  batched   512 repertoires (about 8.6M nucleotide rows, Hamming, eps 0.15,
            minPts 2) in one batched call with the autotuned, graph-captured
            tail; pairs/s is the sum of every repertoire's N(N-1)/2;
  large     one repertoire of 20,000,000 nucleotide CDR3s (24-150 nt,
            100k seeds, 3% substitutions, 0.5% indels), Levenshtein t = 3
            edits + V/J gate, minPts 3: the exact path with the 3-gram bound
            and 5-word Myers (the learned filter reads at most 32 letters).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lanes x clock: 39.3 T ops/s
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def myers_ops_per_letter(W: int) -> int:
    # Xv, Xh (4), Ph (3), Mh, score (4), Ph/Mh shifts (3), Pv (3), Mv = 20 per
    # word and text letter, plus 4 carry/shift ops per extra word
    return 20 * W + 4 * (W - 1)


# absent-letter filter per pair: per direction an AND-NOT of the letter-set
# planes and a bit count, then a maximum and the threshold compare
FILTER_OPS_PER_PAIR = 6
# 3-gram profile bound per pair: 16 four-byte absolute-difference sums, the
# scaled threshold and the compare
QGRAM_OPS_PER_PAIR = 18


def count_kernel_ops(s: dict, vw: int, jw: int) -> int:
    """The count pass's algorithmic work as the work model defines it: the
    V/J gate and length test per gated pair, Myers ops per text letter."""
    W = max(1, (s["max_len"] + 31) // 32)
    gate_ops = 2 * (vw + jw) + 2
    return s["gated_pairs"] * gate_ops + s["myers_letters"] * myers_ops_per_letter(W)


def filter_ops(s: dict) -> int:
    """Ops of the exact lower bound that runs before Myers (absent letters
    or 3-grams), which the work model does not count."""
    if s["band_filter"] == 1:
        return s["gated_pairs"] * FILTER_OPS_PER_PAIR
    if s["band_filter"] == 2:
        return s["gated_pairs"] * QGRAM_OPS_PER_PAIR
    return 0


ENC_FLOPS_PER_ROW = 2 * 32 * (5 * 32 * 128 + 2 * 5 * 128 * 128)  # three conv layers, 32 positions
ENC160_FLOPS_PER_ROW = 5 * ENC_FLOPS_PER_ROW  # the nucleotide encoder: 160 positions
MFMA_BF16_PEAK_TFLOPS = 2500.0


def timed(run, steps: int, warmup: int):
    """ms per call of run() over `steps` calls after `warmup`, bracketed by
    device synchronisation of both runtimes."""
    import torch

    from immunodbscan_amd import hip

    st = None
    for _ in range(warmup):
        st = run()
    hip.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = run()
    hip.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, st


def learned_section(args, cl, exact_rep, exact_ms, exact_res) -> dict:
    from immunodbscan_amd import configs, engine, learned, synth

    t0 = time.perf_counter()
    vocab = learned.vocab_for(synth.LETTERS[synth.AA])
    train = synth.generate(configs.c2_params(n_rows=1_000_000, n_seeds=100_000, seed=41))
    pairs, labels = learned.training_pairs(train, cl, args.train_pairs, seed=3)
    enc = learned.Encoder(vocab, seed=2)
    tp = learned.TrainParams(epochs=args.train_epochs, batch_pairs=4096, lr=5e-4, margin=0.9)
    tstats = enc.train(train, pairs, labels, tp)
    tune = synth.generate(configs.c2_params(n_rows=200_000, n_seeds=20_000, seed=43))
    tau = learned.tune_threshold(enc.embed(tune), learned.neighbour_pairs(tune, cl), recall=0.9998, slack=0.01)
    log(f"learned: trained {tstats['steps']} steps in {tstats['train_ms']:.0f} ms, tau {tau:.4f} "
        f"({time.perf_counter() - t0:.1f} s with labelling)")
    out = {
        "training": {"pairs": int(len(labels)), "positive_fraction": float(labels.mean()), **tstats,
                     "labelled_by": "exact count pass (neighbours) + exact Myers distances (sampled non-neighbours)"},
        "min_similarity": tau,
        "threshold_rule": "0.02% quantile of neighbour-pair similarities on a separate 200k-row repertoire, "
                          "less 0.01 for bfloat16 rounding",
    }

    def one(rep, ex_ms, ex_res, tag):
        lrun = engine.DeviceRun(rep, cl, encoder=enc, min_sim=tau)
        ms, st = timed(lrun, args.steps, args.warmup)
        res = lrun.result()
        s = res.stats
        n = rep.n
        enc_ms = s["encoder_ms"]
        flops = ENC_FLOPS_PER_ROW * n
        return {
            "workload": tag,
            "rows": n,
            "ms_per_step": ms,
            "pairs_per_sec": n * (n - 1) / 2 / (ms * 1e-3),
            "seqs_clustered_per_sec": n / (ms * 1e-3),
            "exact_ms_per_step": ex_ms,
            "speedup_vs_exact": ex_ms / ms,
            "pair_recall": s["neighbour_pairs"] / max(ex_res.stats["neighbour_pairs"], 1),
            "ari": learned.ari(ex_res.clone_id, res.clone_id),
            "proposals": s["gated_pairs"],
            "exact_gated_pairs": ex_res.stats["gated_pairs"],
            "myers_pairs": s["myers_pairs"],
            "exact_myers_pairs": ex_res.stats["myers_pairs"],
            "count_kernel_ms": s["count_kernel_ms"],
            "exact_count_kernel_ms": ex_res.stats["count_kernel_ms"],
            "encoder_ms": enc_ms,
            "encoder_roofline": {"bound": "mfma", "achieved": flops / (enc_ms * 1e-3) / 1e12,
                                 "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": flops / (enc_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS,
                                 "flops_per_row": ENC_FLOPS_PER_ROW},
        }

    if args.learned_c2:
        if exact_rep is None:  # no headline run to compare with: time the exact path here
            exact_rep = synth.generate(configs.c2_params(n_rows=args.rows, n_seeds=max(1, args.rows // 10)))
            erun = engine.DeviceRun(exact_rep, cl)
            exact_ms, _ = timed(erun, args.steps, args.warmup)
            exact_res = erun.result()
        out["c2"] = one(exact_rep, exact_ms, exact_res, "same as the headline line")
        log(f"learned c2: {out['c2']['ms_per_step']:.3f} ms/step, recall {out['c2']['pair_recall']:.5f}")
    if args.c4_rows:
        big = synth.generate(configs.c4_params(n_rows=args.c4_rows))
        erun = engine.DeviceRun(big, cl)
        ems, _ = timed(erun, max(2, args.steps // 4), 1)
        eres = erun.result()
        out["c4"] = one(big, ems, eres, f"{big.n} aa CDR3s, {args.c4_rows // 10} seeds, t=2 edits, minPts=3")
        log(f"learned c4: {out['c4']['ms_per_step']:.3f} ms/step vs exact {ems:.3f}")
    return out


def cpu_literal(reps, cl, threads: int) -> tuple[float, float]:
    """Seconds and all-pairs count of the oracle's literal clustering of each
    repertoire in turn, on `threads` host threads."""
    from oracle import oracle as O

    prm = O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming, nthreads=threads)
    t0 = time.perf_counter()
    pairs = 0.0
    for r in reps:
        O.dbscan(O.Data.of(r), prm, fast=False)
        pairs += r.n * (r.n - 1) / 2
    return time.perf_counter() - t0, pairs


def batched_section(args, threads: int) -> dict:
    from immunodbscan_amd import _lib, configs, engine, synth

    reps = configs.c3_repertoires(n_reps=args.c3_reps)
    cat = synth.concat(reps)
    flags = _lib.IDB_FLAG_AUTOTUNE | _lib.IDB_FLAG_GRAPH
    run = engine.DeviceRun(cat, configs.C3, batch=True, flags=flags)
    ms, st = timed(run, max(3, args.steps // 2), 2)
    s = st.as_dict()
    pairs = float(sum(r.n * (r.n - 1) / 2 for r in reps))
    log(f"batched: {len(reps)} repertoires, {cat.n} rows, {ms:.3f} ms/step")
    # CPU: the literal oracle over the first repertoires, in order, until
    # about 10 s of work; pairs/s extrapolates to the batch
    done, secs, cpairs = 0, 0.0, 0.0
    while not args.no_cpu_baseline and done < len(reps) and secs < 10.0:
        dt, p = cpu_literal(reps[done:done + 8], configs.C3, threads)
        secs += dt
        cpairs += p
        done += min(8, len(reps) - done)
    cpu_v = cpairs / secs if secs else None
    return {
        "workload": f"{len(reps)} nt repertoires, lognormal sizes (median 5k, 200-200k), Hamming eps 0.15, minPts 2",
        "repertoires": len(reps),
        "rows": cat.n,
        "ms_per_step": ms,
        "pairs_per_sec": pairs / (ms * 1e-3),
        "seqs_clustered_per_sec": cat.n / (ms * 1e-3),
        "gated_pairs": s["gated_pairs"],
        "neighbour_pairs": s["neighbour_pairs"],
        "buckets": s["n_buckets"],
        "tiles": s["tiles"],
        "graph_used": s["graph_used"],
        "tuned_launch": {"waves_per_block": s["tuned_wpb"], "grid_x_resident": s["tuned_grid_mul"],
                         "xcd_order": s["tuned_xcd_order"] & 1, "candidates_timed": s["tuned_trials"]},
        "stage_ms": s["stage_ms"],
        "cpu_baseline": None if cpu_v is None else {
            "value": cpu_v, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"literal all-pairs clustering of the first {done} repertoires "
                      f"({sum(r.n for r in reps[:done])} rows) in {secs:.1f} s; extrapolated"},
        "gpu_over_cpu": None if cpu_v is None else (pairs / (ms * 1e-3)) / cpu_v,
    }


def genes_section(args) -> dict:
    """R2 on the device: IMGT-style V calls (1-3 alleles of the 104 reference
    V gene names each) against those names, arrays resident in HBM; the host
    builder on a 200k-row sample beside it."""
    import random

    import numpy as np

    from immunodbscan_amd import _lib, engine, hip
    from immunodbscan_amd.interlink import gene_bitsets, pack_strings

    names = [ln for ln in open(os.path.join(ROOT, "tests", "golden", "VHgenes.txt"), "rb").read().split(b"\n") if ln]
    rng = random.Random(21)
    pool = []
    for _ in range(4096):
        k = rng.randrange(1, 4)
        pool.append(b", or ".join(b"Homsap " + rng.choice(names)[:-1] + b"*0%d F" % rng.randrange(1, 9)
                                  for _ in range(k)))
    pd, po, pl = pack_strings(pool)
    n = args.gene_rows
    pick = np.random.default_rng(22).integers(0, len(pool), n)
    lens = pl[pick].astype(np.uint32)
    off = np.zeros(n, np.uint64)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    data = np.empty(int(lens.sum(dtype=np.uint64)), np.uint8)
    for a in range(0, n, 1 << 20):  # materialise every row's bytes, a million rows at a time
        b = min(n, a + (1 << 20))
        rl = lens[a:b].astype(np.int64)
        base = int(off[a])
        src = np.repeat(po[pick[a:b]].astype(np.int64) - (off[a:b].astype(np.int64) - base), rl)
        data[base:base + int(rl.sum())] = pd[src + np.arange(int(rl.sum()))]
    width = (len(names) + 31) // 32
    gd, go, gl = pack_strings(names)
    ctx = engine.default_context()
    dd, do, dl = (hip.DeviceBuffer.from_array(x) for x in (data, off, lens))
    dout = hip.DeviceBuffer.zeros(4 * n * width)
    call = lambda: ctx.check(_lib.lib.idb_gene_bitsets_gpu(  # noqa: E731
        ctx.handle, _lib.IDB_MEM_DEVICE, n, dd.data_ptr(), do.data_ptr(), dl.data_ptr(), len(names), gd.ctypes.data,
        go.ctypes.data, gl.ctypes.data, dout.data_ptr(), width))
    call()
    hip.synchronize()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        call()
        hip.synchronize()
        ts.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.median(ts))
    # spot check against the host builder on the first rows
    got = dout.to_numpy(np.int32, 20_000 * width).reshape(20_000, width)
    first = [bytes(data[int(off[i]):int(off[i]) + int(lens[i])]) for i in range(200_000)]
    t0 = time.perf_counter()
    host, _ = gene_bitsets(first, names)
    host_s = time.perf_counter() - t0
    if not np.array_equal(got, host[:20_000]):
        raise RuntimeError("device gene bitsets differ from the host builder")
    log(f"genes: {n} calls x {len(names)} names in {ms:.2f} ms")
    return {
        "workload": f"{n} IMGT-style V calls (1-3 alleles each, {data.size / n:.1f} bytes mean) x {len(names)} reference V gene names",
        "rows": n,
        "ms_per_call": ms,
        "calls_per_sec": n / (ms * 1e-3),
        "bytes_per_sec": data.size / (ms * 1e-3),
        "cpu_baseline": {"calls_per_sec": 200_000 / host_s, "cores": 1, "kind": "port",
                         "sample": "first 200,000 calls, host builder (std::string_view::find per name)"},
    }


def pmc_traffic(rows: int):
    """HBM bytes per launch of the count kernel from the committed PMC
    profile of this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by a 256 MiB device copy measured in the same run); None for other sizes."""
    path = os.path.join(ROOT, "profiles", "c2_count_kernel_pmc.json")
    if rows != 200_000 or not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["traffic_bytes"])


def large_section(args, threads: int) -> dict:
    import numpy as np

    from immunodbscan_amd import configs, engine, synth

    rep = synth.generate(configs.c5_params(n_rows=args.c5_rows, n_seeds=max(1, args.c5_rows // 200)))
    log(f"large: generated {rep.n} rows")
    run = engine.DeviceRun(rep, configs.C5)
    ms, st = timed(run, 2, 1)
    s = st.as_dict()
    n = rep.n
    W = max(1, (s["max_len"] + 31) // 32)
    ops = count_kernel_ops(s, rep.vw, rep.jw)
    kms = s["count_kernel_ms"]
    log(f"large: {ms:.1f} ms/step")
    # CPU: the literal oracle on the first rows of the same workload
    sub_n = min(n, args.c5_cpu_rows)
    cpu_v = None
    if not args.no_cpu_baseline:
        dt, cpairs = cpu_literal([rep.subset(np.arange(sub_n))], configs.C5, threads)
        cpu_v = cpairs / dt
    return {
        "workload": f"{n} nt CDR3s, 24-150 nt, t=3 edits + V/J gate, minPts=3",
        "rows": n,
        "ms_per_step": ms,
        "pairs_per_sec": n * (n - 1) / 2 / (ms * 1e-3),
        "seqs_clustered_per_sec": n / (ms * 1e-3),
        "gated_pairs": s["gated_pairs"],
        "myers_pairs": s["myers_pairs"],
        "neighbour_pairs": s["neighbour_pairs"],
        "clusters": s["n_clusters"],
        "myers_words": W,
        "filter": "3-gram profile bound" if s["band_filter"] == 2 else s["band_filter"],
        "stage_ms": s["stage_ms"],
        "count_kernel_roofline": {"bound": "valu", "achieved": ops / (kms * 1e-3) / 1e12, "peak": VALU_PEAK_TOPS,
                                  "unit": "TOP/s", "frac": ops / (kms * 1e-3) / 1e12 / VALU_PEAK_TOPS,
                                  "kernel_ms": kms, "algorithmic_ops": ops, "filter_ops": filter_ops(s)},
        "cpu_baseline": None if cpu_v is None else {
            "value": cpu_v, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"literal all-pairs clustering of the first {sub_n} rows in {dt:.1f} s; extrapolated"},
        "gpu_over_cpu": None if cpu_v is None else (n * (n - 1) / 2 / (ms * 1e-3)) / cpu_v,
        **({"learned": large_learned(args, rep, ms, run.result())} if args.c5_learned else {}),
    }


def large_learned(args, rep, exact_ms, exact_res) -> dict:
    """The learned filter on the same 20M nucleotide rows, side by side with
    the exact path: a 160-position encoder trained on pairs of a separate
    repertoire of the same generator, its threshold set on a third."""
    from immunodbscan_amd import configs, engine, learned, synth

    cl = configs.C5
    t0 = time.perf_counter()
    vocab = learned.vocab_for(synth.LETTERS[synth.NT])
    train = synth.generate(configs.c5_params(n_rows=200_000, n_seeds=1_000, seed=61))
    pairs, labels = learned.training_pairs(train, cl, args.c5_train_pairs, seed=3)
    enc = learned.Encoder(vocab, seed=2, positions=160)
    tp = learned.TrainParams(epochs=args.train_epochs, batch_pairs=4096, lr=5e-4, margin=0.9)
    tstats = enc.train(train, pairs, labels, tp)
    tune = synth.generate(configs.c5_params(n_rows=200_000, n_seeds=1_000, seed=63))
    tau = learned.tune_threshold(enc.embed(tune), learned.neighbour_pairs(tune, cl), recall=0.9998, slack=0.01)
    log(f"large learned: trained {tstats['steps']} steps in {tstats['train_ms']:.0f} ms, tau {tau:.4f} "
        f"({time.perf_counter() - t0:.1f} s with labelling)")
    lrun = engine.DeviceRun(rep, cl, encoder=enc, min_sim=tau)
    ms, _ = timed(lrun, 1, 1)
    res = lrun.result()
    s = res.stats
    enc_ms = s["encoder_ms"]
    flops = ENC160_FLOPS_PER_ROW * rep.n
    log(f"large learned: {ms:.1f} ms/step against exact {exact_ms:.1f}")
    return {
        "training": {"pairs": int(len(labels)), "positive_fraction": float(labels.mean()), **tstats,
                     "repertoire": "200k rows of the same generator (seed 61)"},
        "min_similarity": tau,
        "threshold_rule": "0.02% quantile of neighbour-pair similarities on a separate 200k-row repertoire, "
                          "less 0.01 for bfloat16 rounding",
        "filter_after_proposals": "the exact 3-gram profile bound, then Myers",
        "ms_per_step": ms,
        "exact_ms_per_step": exact_ms,
        "speedup_vs_exact": exact_ms / ms,
        "pair_recall": s["neighbour_pairs"] / max(exact_res.stats["neighbour_pairs"], 1),
        "ari": learned.ari(exact_res.clone_id, res.clone_id),
        "proposals": s["gated_pairs"],
        "exact_gated_pairs": exact_res.stats["gated_pairs"],
        "myers_pairs": s["myers_pairs"],
        "exact_myers_pairs": exact_res.stats["myers_pairs"],
        "stage_ms": s["stage_ms"],
        "exact_stage_ms": exact_res.stats["stage_ms"],
        "count_kernel_ms": s["count_kernel_ms"],
        "exact_count_kernel_ms": exact_res.stats["count_kernel_ms"],
        "encoder_ms": enc_ms,
        "encoder_roofline": {"bound": "mfma", "achieved": flops / (enc_ms * 1e-3) / 1e12,
                             "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": flops / (enc_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS,
                             "flops_per_row": ENC160_FLOPS_PER_ROW},
    }


def cpu_baseline(rep, cl, rows: int, threads: int) -> dict:
    import numpy as np

    from oracle import oracle as O

    sub = rep.subset(np.arange(min(rows, rep.n)))
    d = O.Data.of(sub)
    prm = O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming, nthreads=threads)
    t0 = time.perf_counter()
    O.dbscan(d, prm, fast=False)
    dt = time.perf_counter() - t0
    n = sub.n
    return {
        "value": n * (n - 1) / 2 / dt,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} rows of the workload, full literal all-pairs clustering "
                  f"(duplicates, counts, fuse, labels) in {dt:.2f} s",
        "seqs_per_sec": n / dt,
    }


def host_cores() -> int:
    """Cores this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def c1_section(args, threads: int) -> dict:
    """The 2,000-row Hamming workload (the reference CPU path's own config):
    the GPU call from host memory to labels in host memory, and the oracle's
    literal all-pairs clustering on every host core, median of 3 runs."""
    import numpy as np

    from immunodbscan_amd import configs, engine, synth
    from oracle import oracle as O

    rep = synth.generate(configs.c1_params())
    cl = configs.C1
    run = engine.HostRun(rep, cl)
    ms, _ = timed(run, 50, 5)
    n = rep.n
    d = O.Data.of(rep)
    prm = O.params(cl.eps, cl.abs_k, cl.min_pts, cl.use_hamming, nthreads=threads)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.dbscan(d, prm, fast=False)
        ts.append(time.perf_counter() - t0)
    cpu = float(np.median(ts))
    pairs = n * (n - 1) / 2
    return {
        "workload": f"{n} nt CDR3s, Hamming, eps 0.15 relative, minPts 2 (host memory in and out)",
        "rows": n,
        "ms_per_step": ms,
        "pairs_per_sec": pairs / (ms * 1e-3),
        "seqs_clustered_per_sec": n / (ms * 1e-3),
        "cpu_baseline": {"value": pairs / cpu, "unit": "pairs/s", "cores": threads, "kind": "port",
                         "sample": f"the whole workload, literal all-pairs clustering, median of 3 runs "
                                   f"({', '.join(f'{t * 1e3:.1f}' for t in ts)} ms)"},
        "gpu_over_cpu": (pairs / (ms * 1e-3)) / (pairs / cpu),
    }


def run_sections(args, out, threads, cl, rep, dms, dres) -> None:
    """The side sections, each under its own key of `out`.  They never take
    the headline line down with them: a failure is reported in its place."""
    def section(name, fn):
        try:
            out[name] = fn()
        except Exception as exc:  # noqa: BLE001
            log(f"{name} section failed: {exc!r}")
            out[name] = {"error": repr(exc)}

    if not args.no_c1:
        section("c1", lambda: c1_section(args, threads))
    if not args.no_learned:
        section("learned", lambda: learned_section(args, cl, rep, dms, dres))
    if not args.no_batched:
        section("batched", lambda: batched_section(args, threads))
    if args.c5_rows:
        section("large", lambda: large_section(args, threads))
    if args.gene_rows:
        section("genes", lambda: genes_section(args))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=200_000)
    ap.add_argument("--cpu-rows", type=int, default=200_000)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core this process may use")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip every CPU timing (headline and sections); for profiled runs")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the 200k-row headline and print only the sections run (per-workload profiles)")
    ap.add_argument("--train-pairs", type=int, default=5_000_000)
    ap.add_argument("--train-epochs", type=int, default=4)
    ap.add_argument("--c4-rows", type=int, default=2_000_000, help="0 skips the 2M-row learned comparison")
    ap.add_argument("--no-learned-c2", dest="learned_c2", action="store_false",
                    help="skip the learned filter on the 200k-row workload (the 2M-row line still runs)")
    ap.add_argument("--c3-reps", type=int, default=512)
    ap.add_argument("--c5-rows", type=int, default=20_000_000, help="0 skips the 20M-row nucleotide line")
    ap.add_argument("--c5-cpu-rows", type=int, default=40_000)
    ap.add_argument("--no-c5-learned", dest="c5_learned", action="store_false",
                    help="skip the learned filter on the 20M-row nucleotide line")
    ap.add_argument("--c5-train-pairs", type=int, default=1_000_000)
    ap.add_argument("--gene-rows", type=int, default=10_000_000, help="0 skips the device gene-bitset line")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed steps, run one device-to-device copy of this many MiB "
                         "(the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    # torch's wheel carries its own HIP runtime: initialise it before the
    # library's runtime so that both come up, then use it only to bracket the
    # timed region.
    import torch

    torch.cuda.init()
    torch.cuda.synchronize()
    import numpy as np

    from immunodbscan_amd import configs, engine, hip, synth

    threads = args.cpu_threads or host_cores()
    cl = configs.C2
    if args.no_headline:
        out = {"metric": "sections only (--no-headline)"}
        run_sections(args, out, threads, cl, None, None, None)
        print(json.dumps(out), flush=True)
        return 0
    rep = synth.generate(configs.c2_params(n_rows=args.rows, n_seeds=max(1, args.rows // 10)))
    log(f"generated {rep.n} rows, {len(rep.seq)} letters")
    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        src, dst = hip.DeviceBuffer.zeros(nb), hip.DeviceBuffer(nb)
        hip.copy_device(dst, src, nb)
        hip.synchronize()
        src.free()
        dst.free()
        log(f"copy calibration: {args.copy_calibration_mib} MiB device-to-device")
    # Headline: the boundary's own call, rows in host memory at entry and
    # labels in host memory at exit (H2D copies, clustering, D2H copies).
    run = engine.HostRun(rep, cl)
    for w in range(args.warmup):
        run()
        log(f"warmup {w + 1}/{args.warmup} done")
    hip.synchronize()
    torch.cuda.synchronize()
    cnt_ms, st = [], None
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = run()
        cnt_ms.append(st.count_kernel_ms)
    hip.synchronize()
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    ms = T / args.steps * 1e3
    n = rep.n
    pairs = n * (n - 1) / 2
    s = st.as_dict()
    log(f"timed {args.steps} steps (host memory in and out): {ms:.3f} ms/step; stages {s['stage_ms']}")
    # The same call with rows already in HBM and labels left there.
    drun = engine.DeviceRun(rep, cl)
    dms, _ = timed(drun, args.steps, args.warmup)
    log(f"device-resident: {dms:.3f} ms/step")

    W = max(1, (s["max_len"] + 31) // 32)
    ops = count_kernel_ops(s, rep.vw, rep.jw)
    fops = filter_ops(s)
    kms = float(np.mean(cnt_ms))
    achieved = ops / (kms * 1e-3) / 1e12
    out = {
        "metric": "CDR3 pairs/sec & seqs clustered/sec, exact vs learned filter, 1 MI355X",
        "value": pairs / (ms * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"exact Levenshtein DBSCAN, one repertoire of {n} aa CDR3s (20k-seed generator), "
                        "t=2 edits + V/J gate, minPts=3; rows in host memory at entry, labels in host memory "
                        "at exit",
            "rows": n,
            "mode": "exact",
            "threshold": "abs 2",
            "min_pts": cl.min_pts,
        },
        "seqs_clustered_per_sec": n / (ms * 1e-3),
        "device_resident": {"ms_per_step": dms, "pairs_per_sec": pairs / (dms * 1e-3),
                            "seqs_clustered_per_sec": n / (dms * 1e-3),
                            "note": "rows already in HBM, labels left in HBM"},
        "gated_pairs_per_sec": s["gated_pairs"] / (ms * 1e-3),
        "gated_pairs": s["gated_pairs"],
        "myers_pairs": s["myers_pairs"],
        "neighbour_pairs": s["neighbour_pairs"],
        "clusters": s["n_clusters"],
        "uniques": s["n_uniques"],
        "cores": s["n_cores"],
        "noise": s["n_noise"],
        "stage_ms": s["stage_ms"],
        "roofline": {
            "kernel": ("k_count1<true> (one-word count pass, constant threshold)" if W == 1 and s["band_filter"]
                       else f"k_pairs_myers<{W},false,{1 if s['band_filter'] else 0}> (count pass)"),
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": achieved / VALU_PEAK_TOPS,
            "traffic": pmc_traffic(n),
            "kernel_ms": kms,
            "algorithmic_ops": ops,
            "work_model": "gated pairs x (2(Vw+Jw)+2) + Myers text letters x (20W+4(W-1))",
            "filter_ops": fops,
            "frac_with_filter_ops": (ops + fops) / (kms * 1e-3) / 1e12 / VALU_PEAK_TOPS,
        },
    }
    run_sections(args, out, threads, cl, rep, dms, drun.result())
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle on {min(args.cpu_rows, n)} rows, {threads} threads")
        out["cpu_baseline"] = cpu_baseline(rep, cl, args.cpu_rows, threads)
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
