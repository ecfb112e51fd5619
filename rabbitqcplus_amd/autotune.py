"""Autotuner: picks the correction schedule for a read set on this GPU by timing.

Candidates are the engine and schedule options the library offers:
- staged kernels with one lane per candidate in the alignment, or one wavefront per pair;
- for the lane-per-candidate staged kernels: pipelined batches (no host synchronization
  inside a batch), the same with each batch captured in a HIP graph, or a host
  synchronization per batch;
- for the staged kernels, the MSA's candidate tile (candidates staged into LDS per step):
  its LDS, with the 16-byte-aligned quality rows of long reads, sets how many one-wavefront
  MSA workgroups a compute unit holds;
- the persistent fused kernel at 4, 8 and 12 resident waves per compute unit, its pass
  captured in a HIP graph.
Each candidate corrects the same sample of anchors (after one warm-up pass); the fastest
(median of `reps` passes) wins. All candidates give byte-identical records, so the choice
affects speed only. Choices are cached in a JSON file keyed by the device, its compute
units and the read set's shape (length class, qualities, paired, maps, k).
"""
from __future__ import annotations

import json
import os
import time

from . import engine

CANDIDATES = [
    {"engine": 1, "align_variant": 0},
    {"engine": 1, "align_variant": 0, "use_graphs": 1},
    {"engine": 1, "align_variant": 0, "sync_batches": 1},
    {"engine": 1, "align_variant": 1},
    {"engine": 2, "waves_per_cu": 4, "use_graphs": 1},
    {"engine": 2, "waves_per_cu": 8, "use_graphs": 1},
    {"engine": 2, "waves_per_cu": 12, "use_graphs": 1},
]
# MSA tiles swept for the staged kernels (multiples of 4; the default is 20 with quality
# scores, 64 without)
TILES_Q = (8, 12, 16, 24, 32)
TILES_NOQ = (16, 32, 48)


def candidates(opts) -> list[dict]:
    tiles = TILES_Q if opts.get("use_quality_scores") else TILES_NOQ
    return CANDIDATES + [{"engine": 1, "align_variant": 0, "msa_tile": t} for t in tiles]

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "rabbitqcplus_amd", "autotune.json")


def shape_key(storage, device: int = 0, **opts) -> str:
    name, cus = engine.device_info(device)
    maxlen = int(storage.max_length)
    length_class = 32 * ((maxlen + 31) // 32)
    return (f"{name}|cu{cus}|len{length_class}|q{int(opts.get('use_quality_scores', 0))}|v3|"
            f"pe{int(opts.get('paired', 0))}|maps{int(opts.get('num_hash_maps', 48))}|k{int(opts.get('kmerlength', 0))}")


def _fits(cand, storage) -> bool:
    # the fused kernel keeps the MSA in registers (reads of at most 340 bases)
    return cand.get("engine") != 2 or storage.max_length <= 340


def tune(storage, device: int = 0, sample: int = 65536, reps: int = 3, cache: str | None = DEFAULT_CACHE,
         log=None, **opts) -> dict:
    """Returns {"choice": options, "key": cache key, "timings_ms": [...], "cached": bool}."""
    opts = dict(opts)
    opts.setdefault("paired", int(storage.paired))
    opts.setdefault("use_quality_scores", int(storage.quals is not None))
    key = shape_key(storage, device, **opts)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as fh:
                table = json.load(fh)
        except (OSError, ValueError):
            table = {}
    if key in table:
        return {"choice": table[key]["choice"], "key": key, "timings_ms": table[key]["timings_ms"], "cached": True}
    n = min(sample, storage.n)
    if opts.get("paired"):
        n -= n % 2
    results = []
    for cand in candidates(opts):
        if not _fits(cand, storage):
            continue
        eng = engine.CareEngine(device, **{**opts, **cand})
        try:
            eng.set_reads(storage)
            eng.build_tables()
            eng.correct(0, n)
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                eng.correct(0, n)
                times.append(1e3 * (time.perf_counter() - t0))
        finally:
            eng.close()
        ms = sorted(times)[len(times) // 2]
        results.append({"options": cand, "ms": round(ms, 3)})
        if log:
            log(f"autotune {cand}: {ms:.2f} ms for {n} anchors")
    best = min(results, key=lambda r: r["ms"])
    out = {"choice": best["options"], "key": key, "timings_ms": results, "cached": False}
    if cache:
        table[key] = {"choice": best["options"], "timings_ms": results, "sample": n}
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(table, fh, indent=1)
        os.replace(tmp, cache)
    return out
