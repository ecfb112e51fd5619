#!/usr/bin/env python3
"""Trains the learned seam's classifier on the GPU and writes the classifier file.

Training columns are the decision columns (anchor columns of non-HQ anchors whose base
differs from the consensus) of simulated read sets, labelled by the simulator's truth. The
sets mix the workloads' profiles: uniform genomes with the Q36 -> Q28 profile alone give
few decision columns (a few thousand per 40k reads, with the consensus nearly always
right), so repeat-rich and higher-error genomes, single and paired end, with and without
quality scores, are added. Evaluation sets use other seeds.

    python3 -m rabbitqcplus_amd.train_classifier [--out PATH] [--report PATH] [--epochs N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    __package__ = "rabbitqcplus_amd"

from rabbitqcplus_amd import engine, evaluate, learned, reads, sim  # noqa: E402

# (name, simulator kwargs, use quality scores)
PROFILES = [
    ("uniform_q36", dict(genome_length=200_000, read_length_min=150, read_length_max=150, coverage=30.0,
                         q_first=36.0, q_last=28.0), True),
    ("dup_q30", dict(genome_length=200_000, dup_fraction=0.05, dup_segment_length=1000, dup_copies=2,
                     dup_divergence=0.005, read_length_min=150, read_length_max=150, coverage=30.0,
                     q_first=30.0, q_last=20.0), True),
    ("rep20_q30", dict(genome_length=200_000, repeat_fraction=0.10, repeat_unit_length=2000, repeat_copies=20,
                       repeat_divergence=0.01, read_length_min=150, read_length_max=150, coverage=30.0,
                       q_first=30.0, q_last=20.0), True),
    ("hierr", dict(genome_length=100_000, read_length_min=150, read_length_max=150, coverage=30.0,
                   q_first=24.0, q_last=16.0), True),
    ("hierr_noq", dict(genome_length=100_000, read_length_min=150, read_length_max=150, coverage=30.0,
                       q_first=24.0, q_last=16.0), False),
    ("rep10_hierr", dict(genome_length=150_000, repeat_fraction=0.15, repeat_unit_length=1000, repeat_copies=10,
                         repeat_divergence=0.01, read_length_min=150, read_length_max=150, coverage=30.0,
                         q_first=26.0, q_last=16.0), True),
    ("pe_rep20", dict(genome_length=100_000, repeat_fraction=0.10, repeat_unit_length=2000, repeat_copies=20,
                      repeat_divergence=0.01, read_length_min=150, read_length_max=150, coverage=30.0, paired=1,
                      insert_mean=350.0, insert_sd=50.0, q_first=36.0, q_last=24.0), True),
    ("varlen_rep10", dict(genome_length=150_000, repeat_fraction=0.08, repeat_unit_length=2000, repeat_copies=10,
                          repeat_divergence=0.01, read_length_min=100, read_length_max=250, coverage=30.0,
                          q_first=30.0, q_last=20.0), True),
    ("rep10_noq", dict(genome_length=150_000, repeat_fraction=0.15, repeat_unit_length=1000, repeat_copies=10,
                       repeat_divergence=0.01, read_length_min=150, read_length_max=150, coverage=30.0,
                       q_first=28.0, q_last=18.0), False),
]

TRAIN_SEEDS = (101, 102, 103, 104)
EVAL_SEEDS = (901,)


def dataset(profile, seed):
    name, kw, use_q = profile
    s = sim.simulate(seed=seed, **kw)
    rs = reads.storage_from_sim(s, use_qualities=use_q)
    opts = dict(use_quality_scores=int(use_q), paired=int(rs.paired))
    if rs.max_length >= 160:
        opts["kmerlength"] = 20
    return s, rs, opts


def collect_all(seeds, log):
    Xs, ys = [], []
    for prof in PROFILES:
        for seed in seeds:
            s, rs, opts = dataset(prof, seed)
            X, y, _ = learned.collect(s, rs, **opts)
            Xs.append(X)
            ys.append(y)
            log(f"collect {prof[0]} seed {seed}: {rs.n} reads, {len(y)} decision columns, "
                f"consensus right in {float(np.mean(y)) if len(y) else 0:.3f}")
    return np.concatenate(Xs), np.concatenate(ys)


def evaluate_sets(params, seeds, log):
    rows = []
    for prof in PROFILES:
        for seed in seeds:
            s, rs, opts = dataset(prof, seed)
            _, plain = engine.correct_gpu(rs, **opts)
            _, learn = engine.correct_gpu(rs, classifier=params, **opts)
            cp = evaluate.error_counts(s, rs, plain)
            cl = evaluate.error_counts(s, rs, learn)
            rows.append(dict(profile=prof[0], seed=seed, bases=cp["bases"], before=cp["before"],
                             plain_after=cp["after"], learned_after=cl["after"],
                             plain_miscorrected=cp["miscorrected"], learned_miscorrected=cl["miscorrected"]))
            log(f"eval {prof[0]} seed {seed}: errors {cp['before']} -> plain {cp['after']} / learned {cl['after']} "
                f"(miscorrected {cp['miscorrected']} / {cl['miscorrected']})")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.DEFAULT_CLASSIFIER)
    ap.add_argument("--report", default="")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", flush=True)

    t0 = time.time()
    X, y = collect_all(TRAIN_SEEDS, log)
    log(f"{len(y)} training columns, consensus right in {y.mean():.4f}, {time.time() - t0:.1f} s")
    t1 = time.time()
    params, losses, info = learned.train_classifier(X, y, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                                    seed=args.seed)
    log(f"trained {args.epochs} epochs in {time.time() - t1:.1f} s: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    info.update(profiles=[p[0] for p in PROFILES], train_seeds=list(TRAIN_SEEDS), epoch_losses=losses.tolist())
    learned.save_classifier(args.out, params, info)
    log(f"wrote {args.out}")
    rows = evaluate_sets(params, EVAL_SEEDS, log)
    tot = {k: sum(r[k] for r in rows) for k in ("bases", "before", "plain_after", "learned_after",
                                                  "plain_miscorrected", "learned_miscorrected")}
    log(f"held-out total: {json.dumps(tot)}")
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(dict(info=info, eval=rows, total=tot), fh, indent=1)


if __name__ == "__main__":
    main()
