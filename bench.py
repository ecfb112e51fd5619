#!/usr/bin/env python3
"""Benchmark: corrected reads/s of the CARE correction stage on one MI355X.

A step is one pass of the correction stage over every read of the workload
(candidate lookup -> alignment -> MSA/refinement/decision -> serialized records on
the host), with the read storage and the minhash tables already resident in HBM.
Table construction is timed once and reported separately.

Prints one JSON line (see the keys below). Options:
  --gpus N        only 1 is supported (one GPU per box); N > 1 exits with a message
  --steps K       timed passes (default 3)
  --warmup W      untimed passes (default 1)
  --workload      simulator workload name (default se_5mbp_150bp_30x_dup)
  --no-quality    correct without quality scores
  --cpu-sample S  anchors for the CPU baseline leg (default 300000; 0 disables)
  --cpu-threads T threads for the CPU baseline (default min(16, cores))
  --no-learned    skip the learned leg (the classifier seam with the shipped classifier)
  --no-candidates skip the candidate-correction leg
  --no-autotune   skip the autotuner (which times the schedules on 65,536 anchors first)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from rabbitqcplus_amd import autotune, engine, evaluate, learned, reads, sim  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(t: dict, rs, W: int, M: int):
    """Bytes each kernel must move at minimum, per launch (one launch per batch)."""
    batches = max(1, t["batches"])
    A, C, G = t["anchors"], t["candidates"], t["bucket_values"]
    pitch_b = 4 * rs.pitch_ints
    # k_candidates: length+ambiguity flag, M signatures (8 B) and M table slots (16 B) per
    # anchor, every gathered bucket value (4 B), unique candidate ids written (4 B)
    cand = A * (4 + 1 + 24 * M) + 4 * G + 4 * C
    # k_align (lane per candidate): per candidate its id and length (8 B), fwd+rc HiLo planes
    # (16W B) and the chosen alignment written (16 B); per anchor its fwd planes (8W B), its
    # length and candidate offsets (12 B)
    align = C * (24 + 16 * W) + A * (8 * W + 12)
    # k_msa per candidate: chosen alignment + id (20 B), its 2-bit row once (pitch) and
    # qualities once when enabled; per anchor: its row, offsets, record slot bytes
    qbytes = rs.quals.shape[1] if rs.quals is not None else 0
    msa = C * (20 + pitch_b + qbytes) + A * (pitch_b + 8 + qbytes) + t["record_bytes"]
    # k_fused (one launch per pass) moves the same inputs once, with the candidate lists and
    # alignments kept in wave scratch: per anchor length+flag, signatures, table slots, its
    # planes (8W), row, qualities; every gathered value; per candidate its length (4 B),
    # fwd+rc planes (16W), 2-bit row and qualities; the record bytes
    fused = (A * (4 + 1 + 24 * M + 8 * W + pitch_b + qbytes) + 4 * G
             + C * (4 + 16 * W + pitch_b + qbytes) + t["record_bytes"])
    return {"k_candidates": cand / batches, "k_align": align / batches, "k_msa": msa / batches,
            "k_fused": fused / batches}


def calibration_copy(nbytes: int):
    """One plain device-to-device copy of known size (the runtime's copy kernel), so
    that profiler byte counters can be scaled against it."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for b in (src, dst):
        if hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def cpu_baseline_child(args):
    """CPU-baseline leg in a child process (regenerates the same seeded workload), under
    a time limit and retried once, so that a stalled CPU leg cannot hold the GPU line."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", args.workload,
           "--seed", str(args.seed), "--cpu-sample", str(args.cpu_sample), "--cpu-threads", str(args.cpu_threads)]
    if args.no_quality:
        cmd.append("--no-quality")
    for attempt in range(2):
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=150)
        except subprocess.TimeoutExpired:
            print(f"[bench] CPU baseline attempt {attempt + 1} timed out", file=sys.stderr, flush=True)
            continue
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode == 0 and lines:
            return json.loads(lines[-1])
        print(f"[bench] CPU baseline attempt {attempt + 1} failed: {r.stderr[-500:]}", file=sys.stderr, flush=True)
    return None


def baseline_anchors(sample, rs, paired):
    n = min(sample, rs.n)
    return n - n % 2 if paired else n


def records_digest(data: bytes, below: int):
    """(count, sha256) of the records of reads < below, ordered by read id."""
    import hashlib
    spans = []
    p = 0
    while p < len(data):
        rid = int.from_bytes(data[p:p + 4], "little")
        nb = 8 + (int.from_bytes(data[p + 4:p + 8], "little") & ((1 << 29) - 1))
        if rid < below:
            spans.append((rid, p, nb))
        p += nb
    spans.sort(key=lambda x: x[0])
    return len(spans), hashlib.sha256(b"".join(data[q:q + nb] for _, q, nb in spans)).hexdigest()


def cpu_baseline(rs, opts, sample, threads):
    import oracle
    n = baseline_anchors(sample, rs, opts.get("paired"))
    if oracle.ref_available():
        ref = oracle.CareRef(rs, threads=threads, **opts)
        secs, recs = ref.correct_timed(threads, n)
        # the same anchors once more, keeping the records (sorted by read id), for the
        # comparison with the GPU's
        _, data = ref.correct(threads, n)
        ref.close()
        kind = "reference"
    else:
        cpu = oracle.CareCpu(rs, threads=threads, **opts)
        _, data = cpu.correct(0, n, threads=threads)
        secs = cpu.last_seconds()
        cpu.close()
        kind = "port"
    count, digest = records_digest(data, n)
    return {"value": n / secs, "unit": "reads/s", "cores": threads, "kind": kind,
            "sample": f"first {n} anchors of the same workload (tables over all {rs.n} reads), "
                      f"{secs:.2f} s wall on {threads} threads",
            "records": count, "records_sha256": digest}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the profiler byte counters
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a plain
    copy of known size), as recorded in profiles/traffic.json; None when not recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh).get(workload, {}).get(kernel)
    return None if rec is None else int(rec["bytes_per_launch"])


def timed_passes(eng, steps, warmup, log, tag):
    for w in range(warmup):
        eng.correct()
        log(f"{tag} warmup {w + 1}/{warmup} done")
    step_ms = []
    nrec, data = 0, b""
    for k in range(steps):
        ts = time.perf_counter()
        nrec, data = eng.correct()
        step_ms.append(1e3 * (time.perf_counter() - ts))
        log(f"{tag} step {k + 1}/{steps}: {step_ms[-1]:.1f} ms")
    return float(np.mean(step_ms)), nrec, data, eng.timings()


# What bounds each staged kernel on MI355X (hardware counters, not HBM bandwidth)
LIMITERS = {
    "k_align": "vector and scalar instruction issue (about 1.1 VALU + 0.76 SALU instructions per clock per CU), "
               "not HBM bandwidth",
    "k_msa": "vector issue and LDS/memory latency, not HBM bandwidth",
    "k_candidates": "latency of the random bucket-value loads, not HBM bandwidth",
    "k_fused": "occupancy and instruction issue, not HBM bandwidth",
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="se_5mbp_150bp_30x_dup")
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=300000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-size", type=int, default=0)
    ap.add_argument("--engine", choices=["auto", "staged", "fused"], default="auto",
                    help="auto = staged kernels; fused = one persistent kernel")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-candidates", action="store_true",
                    help="skip the candidate-correction leg (--candidateCorrection)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --engine/--waves-per-cu as given instead of the autotuner's choice")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (profiler byte-counter calibration)")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_baseline_only:
        s = sim.workload(args.workload, seed=args.seed)
        use_q = not args.no_quality
        rs = reads.storage_from_sim(s, use_qualities=use_q)
        copts = dict(use_quality_scores=int(use_q), paired=int(rs.paired))
        if rs.max_length >= 160:
            copts["kmerlength"] = 20
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(json.dumps(cpu_baseline(rs, copts, args.cpu_sample, threads)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the correction stage shards "
              f"by anchor range across GPUs, which is not run here.", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    s = sim.workload(args.workload, seed=args.seed)
    use_q = not args.no_quality
    rs = reads.storage_from_sim(s, use_qualities=use_q)
    opts = dict(use_quality_scores=int(use_q), paired=int(rs.paired), batch_size=args.batch_size,
                engine={"auto": 0, "staged": 1, "fused": 2}[args.engine], waves_per_cu=args.waves_per_cu)
    if rs.max_length >= 160:
        opts["kmerlength"] = 20
    log(f"workload {args.workload}: {rs.n} reads, max length {rs.max_length}, simulated in {time.time() - t0:.1f} s")
    tuned = None
    if not args.no_autotune and args.engine == "auto":
        base = {k: v for k, v in opts.items() if k not in ("engine", "waves_per_cu")}
        tuned = autotune.tune(rs, log=log, cache=None, **base)
        opts.update(tuned["choice"])
        log(f"autotune choice {tuned['choice']}")

    eng = engine.CareEngine(0, **opts)
    if args.calib_copy_mb > 0:
        calibration_copy(args.calib_copy_mb << 20)
        log(f"calibration copy of {args.calib_copy_mb} MiB done")
    t1 = time.time()
    eng.set_reads(rs)
    t2 = time.time()
    maps = eng.build_tables()
    t3 = time.time()
    log(f"reads to device {1e3 * (t2 - t1):.1f} ms, tables ({maps} maps) {1e3 * (t3 - t2):.1f} ms")
    ms_per_step, nrec, data, t = timed_passes(eng, args.steps, args.warmup, log, "plain")
    eng.close()
    value = rs.n / (ms_per_step / 1e3)

    W = max(1, (rs.max_length + 31) // 32)
    per_launch = algorithmic_bytes(t, rs, W, eng.opts.num_hash_maps)
    batches = max(1, t["batches"])
    if t["fused_kernel_ms"] > 0:
        kernel_ms = {"k_fused": t["fused_kernel_ms"] / batches}
    else:
        kernel_ms = {"k_candidates": t["cand_kernel_ms"] / batches, "k_align": t["align_kernel_ms"] / batches,
                     "k_msa": t["msa_kernel_ms"] / batches}
    dominant = max(kernel_ms, key=kernel_ms.get)
    achieved = per_launch[dominant] / (kernel_ms[dominant] / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.workload, dominant),
                "kernel": dominant, "avg_launch_ms": round(kernel_ms[dominant], 4),
                "algorithmic_bytes_per_launch": int(per_launch[dominant]), "launches_per_step": int(batches),
                "limiter": LIMITERS.get(dominant, "instruction issue and latency, not HBM bandwidth") +
                           " (counters: profiles/pmc_staged_se_5mbp_150bp_1M.md)"}
    plain_err = evaluate.error_counts(s, rs, data)
    log(f"GPU plain: {value:.0f} reads/s, {ms_per_step:.1f} ms/pass; kernel avg ms {json.dumps(kernel_ms)}; "
        f"errors {plain_err}")

    # FASTQ text -> device read storage, and records -> corrected FASTQ text (the stages
    # either side of the correction pass; each timed once)
    texts = sim.fastq_texts(s)
    text_bytes = sum(len(t) for t in texts)
    eng = engine.CareEngine(0, **opts)
    ta = time.time()
    eng.set_reads_fastq(*texts)
    tb = time.time()
    outs = [eng.write_fastq(data, f) for f in range(len(texts))]
    tc = time.time()
    same_storage = bool(np.array_equal(eng.inspect_reads()["enc"], rs.enc))
    eng.write_fastq(data, 0)
    tio = eng.timings()
    eng.close()
    out_bytes = sum(len(o) for o in outs)
    io_line = {
        "fastq_bytes": text_bytes, "storage_equal_to_array_path": same_storage,
        "parse_device_ms": round(tio["fastq_device_ms"], 3), "parse_wall_ms": round(1e3 * (tb - ta), 2),
        "parse_device_gbps": round(text_bytes / (tio["fastq_device_ms"] / 1e3) / 1e9, 2),
        "merge_output_bytes": out_bytes, "merge_device_ms_file0": round(tio["merge_device_ms"], 3),
        "merge_wall_ms_all_files": round(1e3 * (tc - tb), 2),
    }
    log(f"FASTQ parse {text_bytes / 1e6:.1f} MB: device {tio['fastq_device_ms']:.2f} ms, wall {1e3 * (tb - ta):.1f} ms; "
        f"merge device {tio['merge_device_ms']:.2f} ms (file 0), wall {1e3 * (tc - tb):.1f} ms")
    del texts, outs

    learned_line = None
    if not args.no_learned and os.path.exists(learned.DEFAULT_CLASSIFIER) and rs.max_length <= 340:
        params, _ = learned.load_classifier()
        eng = engine.CareEngine(0, correction_type=1, **opts)
        eng.set_reads(rs)
        eng.build_tables()
        eng.set_classifier(params)
        lms, lrec, ldata, lt = timed_passes(eng, args.steps, args.warmup, log, "learned")
        eng.close()
        lerr = evaluate.error_counts(s, rs, ldata)
        lvalue = rs.n / (lms / 1e3)
        learned_line = {"value": round(lvalue, 1), "unit": "reads/s", "ms_per_step": round(lms, 3),
                        "vs_plain": round(lvalue / value, 4), "records": lrec,
                        "error_rate_after": lerr["after"] / lerr["bases"], "errors_after": lerr["after"],
                        "fixed": lerr["fixed"], "miscorrected": lerr["miscorrected"],
                        "classifier": os.path.relpath(learned.DEFAULT_CLASSIFIER, ROOT)}
        log(f"GPU learned: {lvalue:.0f} reads/s, {lms:.1f} ms/pass; errors {lerr}")

    # candidate correction (opt-in in the reference): the same passes with HQ anchors also
    # correcting their candidates; records then hold several results per read
    cand_line = None
    if not args.no_candidates and rs.max_length <= 340:
        copts = {k: v for k, v in opts.items() if k not in ("engine", "waves_per_cu", "use_graphs")}
        eng = engine.CareEngine(0, correct_candidates=1, engine=1, **copts)
        eng.set_reads(rs)
        eng.build_tables()
        cms, crec, cdata, ct = timed_passes(eng, args.steps, args.warmup, log, "candidates")
        eng.close()
        cvalue = rs.n / (cms / 1e3)
        cand_line = {"value": round(cvalue, 1), "unit": "reads/s", "ms_per_step": round(cms, 3),
                     "vs_plain": round(cvalue / value, 4), "records": crec, "record_bytes": len(cdata)}
        log(f"GPU candidates: {cvalue:.0f} reads/s, {cms:.1f} ms/pass, {crec} records")
        del cdata

    cpu = None
    if args.cpu_sample > 0:
        log(f"CPU baseline, {args.cpu_sample} anchors")
        cpu = cpu_baseline_child(args)
        if cpu is not None and "records_sha256" in cpu:
            # the plain pass's records of the same anchors against the CPU run's
            n = baseline_anchors(args.cpu_sample, rs, rs.paired)
            gcount, gdigest = records_digest(data, n)
            cpu["records_identical_to_gpu"] = bool(gcount == cpu["records"] and gdigest == cpu["records_sha256"])
            log(f"CPU baseline records of the first {n} anchors: {cpu['records']}, identical to the GPU's: "
                f"{cpu['records_identical_to_gpu']}")

    line = {
        "metric": "corrected reads/s on 1 MI355X, plain vs learned; post-correction error rate",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (seeded simulator)",
        "config": {"workload": args.workload, "reads": rs.n, "read_length": [int(rs.lengths.min()), rs.max_length],
                   "coverage": 30, "k": int(eng.opts.kmerlength) or (20 if rs.max_length < 160 else 32),
                   "maps": int(eng.opts.num_hash_maps), "quality_scores": use_q, "paired": bool(rs.paired),
                   "path": "plain", "engine": "fused" if "k_fused" in kernel_ms else "staged",
                   "schedule": {k: opts[k] for k in ("engine", "waves_per_cu", "align_variant", "use_graphs") if k in opts}},
        "autotune": tuned["timings_ms"] if tuned else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "error_rate_before": plain_err["before"] / plain_err["bases"],
        "error_rate_after": plain_err["after"] / plain_err["bases"],
        "errors": plain_err,
        "learned": learned_line,
        "candidate_correction": cand_line,
        "records": nrec,
        "fastq_io": io_line,
        "table_build_ms": round(1e3 * (t3 - t2), 2),
        "stage_ms": {"candidates": round(t["candidates_ms"], 3), "align": round(t["align_ms"], 3),
                     "msa": round(t["msa_ms"], 3), "output": round(t["output_ms"], 3)},
        "kernel_avg_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "pairs_aligned": int(t["candidates"]),
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
