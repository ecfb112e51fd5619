#!/usr/bin/env python3
"""Benchmark: corrected reads/s of the CARE correction stage on one MI355X.

A step is one pass of the correction stage over every read of the workload
(candidate lookup -> alignment -> MSA/refinement/decision -> serialized records on
the host), with the read storage and the minhash tables already resident in HBM.
Table construction is timed once and reported separately, and so is one end-to-end
run (reads uploaded, tables built, one pass).

The default workload is the largest one: 8.57M single-end reads of 100..250 bp at 30x
from a 50 Mbp genome with 10-copy repeat families (k = 20, 48 maps, qualities on).

Prints one JSON line. Legs beside the headline (plain path, bit-exact to the reference):
  learned               the learned form: the MLP classifier seam on seeded alignment
  learned_exact         the classifier seam on the reference's shift scan
  pe_workload           2.0M read pairs with 20-copy repeats: its own autotuner sweep, plain
                        and learned
  fastq_io              FASTQ text -> device read storage, records -> corrected FASTQ text
  candidate_correction  --candidateCorrection passes
  cpu_baseline          the reference engine compiled from source, on every host core,
                        over a bounded sample of the same workload's anchors
Options:
  --gpus N          only 1 is supported (one GPU per box); N > 1 exits with a message
  --steps K         timed passes (default 3)
  --warmup W        untimed passes (default 1)
  --workload        simulator workload name (default se_50mbp_100to250bp_30x_rep)
  --no-quality      correct without quality scores
  --cpu-sample S    anchors for the CPU baseline leg (default 300000; 0 disables)
  --cpu-threads T   threads for the CPU baseline (default: every core this process may use)
  --pe-workload     paired workload of the PE leg ('' disables)
  --no-learned / --no-candidates / --no-fastq / --no-autotune   skip those legs
  --engine / --waves-per-cu / --batch-size / --graphs   a fixed schedule (with --no-autotune)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from rabbitqcplus_amd import autotune, engine, evaluate, learned, reads, sim  # noqa: E402

HBM_PEAK_GBS = 8000.0
MAPS_SLOT_BYTES = 16      # one 16-byte AoS table slot per map and anchor


def kmer_length(rs):
    """k = 20 for every workload here (the reference picks 32 for reads of 160 bp and
    more, src/dispatch_care_correct_cpu.cpp:283-297; the workloads fix k = 20)."""
    return 20


def algorithmic_bytes(t: dict, rs, use_q: bool, M: int):
    """Algorithmic bytes of the correction stage, by the per-anchor terms of the path:
    own 2-bit row + length (4*pitch + 4), one 16-byte table slot per map, every gathered
    bucket value (4 B), every candidate's row + length (4*pitch + 4), qualities of the
    anchor and its candidates (L each; the candidates kept after the filters are at most
    the aligned ones, which are counted), and the serialized records. Returns the bytes
    of one pass split by the kernel that moves each term, and the pass total."""
    A, C, G = t["anchors"], t["candidates"], t["bucket_values"]
    row = 4 * rs.pitch_ints + 4
    mean_len = float(rs.lengths.mean())
    cand = A * M * MAPS_SLOT_BYTES + 4 * G
    align = row * (A + C)
    msa = (mean_len * (A + C) if use_q else 0.0) + t["record_bytes"]
    return {"k_candidates": cand, "k_align": align, "k_msa": msa, "pass": cand + align + msa}


def host_info(threads):
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "threads_used": threads}
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, p = fh.read().split()
            info["cgroup_cpus"] = None if q == "max" else round(int(q) / int(p), 2)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                info["cpu_model"] = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def default_threads():
    """Every core this process may run on (its affinity set, bounded by a cgroup quota)."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, p = fh.read().split()
            if q != "max":
                n = min(n, max(1, int(int(q) // int(p))))
    except (OSError, ValueError):
        pass
    return n


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


def cpu_baseline(rs, opts, sample, threads, tables):
    """The reference engine compiled from source (kind "reference"), its tables read by its
    own loader from the engine's table file (the reference's format), timed over the
    first `sample` anchors on `threads` threads; the restatement's parallel table build on
    the same threads is timed beside it (the reference builds its tables only sequentially
    without its thread pool)."""
    import oracle
    n = baseline_anchors(sample, rs, opts.get("paired"))
    out = {"unit": "reads/s", "cores": threads}
    t0 = time.time()
    cpu = oracle.CareCpu(rs, threads=threads, **opts)
    out["table_build_s"] = round(time.time() - t0, 2)
    out["table_build_kind"] = "port (the restatement's tables equal the reference's)"
    cpu.close()
    del cpu
    if oracle.ref_available():
        ref = oracle.CareRef(rs, threads=threads, tables=tables, **opts)
        out["tables_loaded_s"] = round(ref.load_seconds, 2)
        secs, _ = ref.correct_timed(threads, n)
        # the same anchors once more, keeping the records (sorted by read id), for the
        # comparison with the GPU's
        _, data = ref.correct(threads, n)
        ref.close()
        out["kind"] = "reference"
    else:
        cpu = oracle.CareCpu(rs, threads=threads, **opts)
        _, data = cpu.correct(0, n, threads=threads)
        secs = cpu.last_seconds()
        cpu.close()
        out["kind"] = "port"
    count, digest = records_digest(data, n)
    out.update({"value": round(n / secs, 1), "seconds": round(secs, 3),
                "sample": f"first {n} anchors of the same workload (tables over all {rs.n} reads), "
                          f"{secs:.2f} s wall on {threads} threads",
                "records": count, "records_sha256": digest})
    return out


def cpu_baseline_child(args, tables, threads, log):
    """CPU-baseline leg in a child process (it regenerates the same seeded workload) under
    a time limit; a failure or time-out is recorded, not retried."""
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", args.workload,
           "--seed", str(args.seed), "--cpu-sample", str(args.cpu_sample), "--cpu-threads", str(threads),
           "--tables", tables]
    if args.no_quality:
        cmd.append("--no-quality")
    t0 = time.time()
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=args.cpu_timeout)
    except subprocess.TimeoutExpired:
        log(f"CPU baseline timed out after {args.cpu_timeout} s")
        return {"error": f"timed out after {args.cpu_timeout} s (not retried)"}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        log(f"CPU baseline failed: {r.stderr[-800:]}")
        return {"error": f"child exited with {r.returncode}: {r.stderr[-300:]}"}
    res = json.loads(lines[-1])
    res["child_wall_s"] = round(time.time() - t0, 1)
    return res


def measured(workload, kernel, what="traffic"):
    """Profiler figures recorded in profiles/traffic.json for `kernel` on `workload`."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh).get(workload, {}).get(kernel)
    if rec is None:
        return None
    return int(rec["bytes_per_launch"]) if what == "traffic" else rec


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


def engine_opts(rs, use_q, args):
    o = dict(use_quality_scores=int(use_q), paired=int(rs.paired), batch_size=args.batch_size,
             engine={"auto": 0, "staged": 1, "fused": 2}[args.engine], waves_per_cu=args.waves_per_cu,
             kmerlength=kmer_length(rs))
    if args.graphs:
        o["use_graphs"] = 1
    return o


def run_leg(rs, opts, steps, warmup, log, tag, classifier=None):
    eng = engine.CareEngine(0, **opts)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        if classifier is not None:
            eng.set_classifier(classifier)
        return timed_passes(eng, steps, warmup, log, tag)
    finally:
        eng.close()


def error_line(s, rs, data):
    e = evaluate.error_counts(s, rs, data)
    return {"errors_before": e["before"], "errors_after": e["after"], "bases": e["bases"],
            "error_rate_before": e["before"] / e["bases"], "error_rate_after": e["after"] / e["bases"],
            "fixed": e["fixed"], "miscorrected": e["miscorrected"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="se_50mbp_100to250bp_30x_rep")
    ap.add_argument("--pe-workload", default="pe_20mbp_2x150bp_30x_rep")
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=300000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-timeout", type=int, default=600)
    ap.add_argument("--batch-size", type=int, default=0)
    ap.add_argument("--engine", choices=["auto", "staged", "fused"], default="auto",
                    help="auto = the autotuner's choice; staged = per-stage kernels; fused = one persistent kernel")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--graphs", action="store_true",
                    help="capture each pass's batches in HIP graphs (with --no-autotune or --engine)")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-candidates", action="store_true",
                    help="skip the candidate-correction leg (--candidateCorrection)")
    ap.add_argument("--no-fastq", action="store_true", help="skip the FASTQ parse / merge leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --engine/--waves-per-cu as given instead of the autotuner's choice")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (profiler byte-counter calibration)")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tables", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_baseline_only:
        s = sim.workload(args.workload, seed=args.seed)
        use_q = not args.no_quality
        rs = reads.storage_from_sim(s, use_qualities=use_q)
        del s
        copts = dict(use_quality_scores=int(use_q), paired=int(rs.paired), kmerlength=kmer_length(rs))
        threads = args.cpu_threads or default_threads()
        print(json.dumps(cpu_baseline(rs, copts, args.cpu_sample, threads, args.tables)), flush=True)
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
    opts = engine_opts(rs, use_q, args)
    log(f"workload {args.workload}: {rs.n} reads, lengths {int(rs.lengths.min())}..{rs.max_length}, "
        f"simulated in {time.time() - t0:.1f} s")
    # end to end once, in the process's first context (a RabbitQCPlus run corrects once per
    # process): reads to the device, tables (signatures, sort-based build, every read's own
    # bucket), one pass, with the library's default schedule. Later contexts of the same
    # process can pay more for their first large allocation (device memory that another
    # context freed is prepared again); their timed passes follow warm-up passes.
    eng = engine.CareEngine(0, **opts)
    t1 = time.perf_counter()
    eng.set_reads(rs)
    t2 = time.perf_counter()
    eng.build_tables()
    t3 = time.perf_counter()
    eng.correct()
    t4 = time.perf_counter()
    eng.close()
    e2e = {"upload_ms": round(1e3 * (t2 - t1), 2), "tables_ms": round(1e3 * (t3 - t2), 2),
           "first_pass_ms": round(1e3 * (t4 - t3), 2), "total_ms": round(1e3 * (t4 - t1), 2),
           "reads_per_s": round(rs.n / (t4 - t1), 1), "schedule": "default"}
    log(f"end to end: {json.dumps(e2e)}")

    tuned = None
    if not args.no_autotune and args.engine == "auto":
        base = {k: v for k, v in opts.items() if k not in ("engine", "waves_per_cu")}
        # a sample of 2^19 anchors: schedules that differ by under 1% on 2^16 anchors (MSA
        # tiles) can differ by 4% over a whole pass, so the sample must reach batch scale
        tuned = autotune.tune(rs, sample=1 << 19, log=log, cache=None, **base)
        opts.update(tuned["choice"])
        log(f"autotune choice {tuned['choice']}")

    if args.calib_copy_mb > 0:
        calibration_copy(args.calib_copy_mb << 20)
        log(f"calibration copy of {args.calib_copy_mb} MiB done")

    eng = engine.CareEngine(0, **opts)
    eng.set_reads(rs)
    maps = eng.build_tables()
    ms_per_step, nrec, data, t = timed_passes(eng, args.steps, args.warmup, log, "plain")
    M = int(eng.opts.num_hash_maps)
    cpu_tables = None
    if args.cpu_sample > 0:
        fd, cpu_tables = tempfile.mkstemp(prefix="care_tables_", suffix=".bin")
        os.close(fd)
        eng.save_tables(cpu_tables)
    eng.close()
    value = rs.n / (ms_per_step / 1e3)

    if t.get("graph_ms", 0) > 0 and t["align_kernel_ms"] == 0:
        # captured batches carry no stage events: the same kernels, uncaptured, give the
        # per-kernel durations (one untimed pass after a warm-up)
        _, _, _, t_stages = run_leg(rs, dict(opts, use_graphs=0), 1, 1, log, "stage timings")
        for key in ("cand_kernel_ms", "align_kernel_ms", "msa_kernel_ms", "batches"):
            t[key] = t_stages[key]
    batches = max(1, t["batches"])
    if t["fused_kernel_ms"] > 0:
        kernel_ms = {"k_fused": t["fused_kernel_ms"] / batches}
    else:
        kernel_ms = {"k_candidates": t["cand_kernel_ms"] / batches, "k_align": t["align_kernel_ms"] / batches,
                     "k_msa": t["msa_kernel_ms"] / batches}
    ab = algorithmic_bytes(t, rs, use_q, M)
    dominant = max(kernel_ms, key=kernel_ms.get)
    per_launch = (ab["pass"] if dominant == "k_fused" else ab[dominant]) / batches
    achieved = per_launch / (kernel_ms[dominant] / 1e3) / 1e9
    counters = measured(args.workload, dominant, "record") or {}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured(args.workload, dominant),
                "kernel": dominant, "avg_launch_ms": round(kernel_ms[dominant], 4),
                "algorithmic_bytes_per_launch": int(per_launch), "launches_per_step": int(batches),
                "bytes_model": "per anchor: own row+length, 48 x 16 B slots, 4 B per bucket value, row+length per "
                               "candidate, qualities of anchor and candidates, record bytes (split by kernel)",
                "issue": counters.get("issue"), "counters_source": counters.get("source")}
    pass_frac = ab["pass"] / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS
    path_roofline = {"algorithmic_bytes_per_step": int(ab["pass"]),
                     "achieved": round(ab["pass"] / (ms_per_step / 1e3) / 1e9, 2), "unit": "GB/s",
                     "frac": round(pass_frac, 5)}
    plain_err = error_line(s, rs, data)
    log(f"GPU plain: {value:.0f} reads/s, {ms_per_step:.1f} ms/pass; kernel avg ms {json.dumps(kernel_ms)}; "
        f"errors {plain_err['errors_after']}")

    learned_line = exact_line = None
    params = None
    if not args.no_learned and os.path.exists(learned.DEFAULT_CLASSIFIER) and rs.max_length <= 340:
        params, _ = learned.load_classifier()
        lo = dict(opts, correction_type=1, seeded_alignment=1, engine=1)
        lo.pop("use_graphs", None)
        lms, lrec, ldata, lt = run_leg(rs, lo, args.steps, args.warmup, log, "learned", params)
        lvalue = rs.n / (lms / 1e3)
        learned_line = {"value": round(lvalue, 1), "unit": "reads/s", "ms_per_step": round(lms, 3),
                        "vs_plain": round(lvalue / value, 4), "records": lrec,
                        "form": "MLP classifier seam (bf16, fallback below confidence 0.9) on seeded alignment",
                        "kernel_avg_ms": {"k_candidates": round(lt["cand_kernel_ms"] / max(1, lt["batches"]), 4),
                                          "k_align_seeded": round(lt["align_kernel_ms"] / max(1, lt["batches"]), 4),
                                          "k_msa": round(lt["msa_kernel_ms"] / max(1, lt["batches"]), 4)},
                        "classifier": os.path.relpath(learned.DEFAULT_CLASSIFIER, ROOT)}
        learned_line.update(error_line(s, rs, ldata))
        log(f"GPU learned: {lvalue:.0f} reads/s ({lvalue / value:.3f}x plain), errors {learned_line['errors_after']}")
        eo = dict(opts, correction_type=1)
        ems, erec, edata, _ = run_leg(rs, eo, args.steps, args.warmup, log, "learned_exact", params)
        evalue = rs.n / (ems / 1e3)
        exact_line = {"value": round(evalue, 1), "ms_per_step": round(ems, 3), "vs_plain": round(evalue / value, 4),
                      "form": "MLP classifier seam on the reference's shift scan"}
        exact_line.update(error_line(s, rs, edata))
        del ldata, edata

    # candidate correction (opt-in in the reference): the same passes with HQ anchors also
    # correcting their candidates; records then hold several results per read
    cand_line = None
    if not args.no_candidates and rs.max_length <= 340:
        copts = {k: v for k, v in opts.items() if k not in ("engine", "waves_per_cu", "use_graphs")}
        cms, crec, cdata, ct = run_leg(rs, dict(copts, correct_candidates=1, engine=1), args.steps, args.warmup,
                                       log, "candidates")
        cvalue = rs.n / (cms / 1e3)
        cand_line = {"value": round(cvalue, 1), "unit": "reads/s", "ms_per_step": round(cms, 3),
                     "vs_plain": round(cvalue / value, 4), "records": crec, "record_bytes": len(cdata)}
        del cdata

    # FASTQ text -> device read storage, and records -> corrected FASTQ text (the stages
    # either side of the correction pass; each timed once)
    io_line = None
    if not args.no_fastq:
        texts = sim.fastq_texts(s)
        text_bytes = int(sum(len(x) for x in texts))
        eng = engine.CareEngine(0, **opts)
        try:
            ta = time.perf_counter()
            eng.set_reads_fastq(*texts)
            tb = time.perf_counter()
            outs = [eng.write_fastq(data, f, as_array=True) for f in range(len(texts))]
            tc = time.perf_counter()
            same_storage = bool(np.array_equal(eng.inspect_reads()["enc"], rs.enc))
            tio = eng.timings()
        finally:
            eng.close()
        out_bytes = sum(len(o) for o in outs)
        io_line = {
            "fastq_bytes": text_bytes, "storage_equal_to_array_path": same_storage,
            "parse_device_ms": round(tio["fastq_device_ms"], 3), "parse_wall_ms": round(1e3 * (tb - ta), 2),
            "parse_device_gbps": round(text_bytes / (tio["fastq_device_ms"] / 1e3) / 1e9, 2),
            "merge_output_bytes": out_bytes, "merge_device_ms_last_file": round(tio["merge_device_ms"], 3),
            "merge_wall_ms_all_files": round(1e3 * (tc - tb), 2),
        }
        log(f"FASTQ parse {text_bytes / 1e6:.1f} MB: device {tio['fastq_device_ms']:.2f} ms; "
            f"merge wall {1e3 * (tc - tb):.1f} ms")
        del texts, outs

    # the paired workload: plain and learned passes
    pe_line = None
    if args.pe_workload:
        sp = sim.workload(args.pe_workload, seed=args.seed)
        rp = reads.storage_from_sim(sp, use_qualities=use_q)
        po = dict(use_quality_scores=int(use_q), paired=1, kmerlength=kmer_length(rp), engine=1)
        ptuned = None
        if not args.no_autotune and args.engine == "auto":
            # the paired workload's own sweep (its shape key differs: pairs, 150 bp, k)
            ptuned = autotune.tune(rp, sample=1 << 19, log=log, cache=None,
                                   **{k: v for k, v in po.items() if k != "engine"})
            po.update(ptuned["choice"])
            log(f"pe autotune choice {ptuned['choice']}")
        pms, prec, pdata, pt = run_leg(rp, po, args.steps, args.warmup, log, "pe plain")
        pe_line = {"workload": args.pe_workload, "reads": rp.n,
                   "schedule": {k: po[k] for k in ("engine", "waves_per_cu", "align_variant", "msa_tile", "use_graphs",
                                                   "sync_batches") if k in po},
                   "autotune": ptuned["timings_ms"] if ptuned else None, "plain": {
            "value": round(rp.n / (pms / 1e3), 1), "ms_per_step": round(pms, 3), "records": prec,
            "pairs_aligned": int(pt["candidates"]), "bucket_values": int(pt["bucket_values"])}}
        pe_line["plain"].update(error_line(sp, rp, pdata))
        if params is not None:
            lms, _, ldata, _ = run_leg(rp, dict(po, correction_type=1, seeded_alignment=1), args.steps, args.warmup,
                                       log, "pe learned", params)
            pe_line["learned"] = {"value": round(rp.n / (lms / 1e3), 1), "ms_per_step": round(lms, 3),
                                  "vs_plain": round(pms / lms, 4)}
            pe_line["learned"].update(error_line(sp, rp, ldata))
        del sp, rp

    cpu = None
    threads = args.cpu_threads or default_threads()
    if args.cpu_sample > 0:
        log(f"CPU baseline, {args.cpu_sample} anchors on {threads} threads")
        try:
            cpu = cpu_baseline_child(args, cpu_tables, threads, log)
        finally:
            os.unlink(cpu_tables)
        cpu.update(host_info(threads))
        if "records_sha256" in cpu:
            # the plain pass's records of the same anchors against the CPU run's
            n = baseline_anchors(args.cpu_sample, rs, rs.paired)
            gcount, gdigest = records_digest(data, n)
            cpu["records_identical_to_gpu"] = bool(gcount == cpu["records"] and gdigest == cpu["records_sha256"])
            log(f"CPU baseline {cpu['value']:.0f} reads/s; records of the first {n} anchors identical to the GPU's: "
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
                   "coverage": 30, "k": kmer_length(rs), "maps": M, "quality_scores": use_q, "paired": bool(rs.paired),
                   "path": "plain (bit-exact to the reference)",
                   "engine": "fused" if "k_fused" in kernel_ms else "staged",
                   "schedule": {k: opts[k] for k in ("engine", "waves_per_cu", "align_variant", "use_graphs", "sync_batches",
                                                         "msa_tile")
                                if k in opts}},
        "learned": learned_line,
        "learned_exact": exact_line,
        "roofline": roofline,
        "path_roofline": path_roofline,
        "cpu_baseline": cpu,
        "end_to_end": e2e,
        "table_build_ms": e2e["tables_ms"],
        "error_rate_before": plain_err["error_rate_before"],
        "error_rate_after": plain_err["error_rate_after"],
        "errors": plain_err,
        "pe_workload": pe_line,
        "candidate_correction": cand_line,
        "fastq_io": io_line,
        "autotune": tuned["timings_ms"] if tuned else None,
        "records": nrec,
        "stage_ms": {"candidates": round(t["candidates_ms"], 3), "align": round(t["align_ms"], 3),
                     "msa": round(t["msa_ms"], 3), "output": round(t["output_ms"], 3)},
        "kernel_avg_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "pairs_aligned": int(t["candidates"]),
        "bucket_values": int(t["bucket_values"]),
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu and "value" in cpu else None,
        "gpu_end_to_end_vs_cpu": round(e2e["reads_per_s"] / cpu["value"], 1) if cpu and "value" in cpu else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
