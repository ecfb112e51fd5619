"""Shared workload builders for the parity tests (seeded, small enough for the oracle)."""
from __future__ import annotations

import functools

import numpy as np

from rabbitqcplus_amd import reads, sim

# name -> (simulator kwargs, engine/oracle option overrides)
CASES = {
    # plain SE at the 200 kbp / 150 bp / 30x size, qualities off (RabbitQCPlus default)
    "se": (dict(genome_length=200_000, read_length_min=150, read_length_max=150, coverage=30.0, seed=11), {}),
    # same reads, --useQualityScores
    "se_q": (dict(genome_length=200_000, read_length_min=150, read_length_max=150, coverage=30.0, seed=11),
             dict(use_quality_scores=1)),
    # high error profile (Q24 -> Q16, ~1.5% errors) so that many reads are corrected
    "se_higherr": (dict(genome_length=60_000, read_length_min=150, read_length_max=150, coverage=30.0,
                        q_first=24.0, q_last=16.0, seed=12), {}),
    "se_higherr_q": (dict(genome_length=60_000, read_length_min=150, read_length_max=150, coverage=30.0,
                          q_first=24.0, q_last=16.0, seed=12), dict(use_quality_scores=1)),
    # repeats and duplicated segments: heavy-tailed candidate counts, refinement rounds
    "se_repeats": (dict(genome_length=120_000, repeat_fraction=0.2, repeat_unit_length=1000, repeat_copies=10,
                        repeat_divergence=0.01, dup_fraction=0.05, dup_segment_length=800, dup_copies=2,
                        dup_divergence=0.005, read_length_min=150, read_length_max=150, coverage=30.0,
                        q_first=30.0, q_last=20.0, seed=13), {}),
    # paired end, insert N(350,50)
    "pe": (dict(genome_length=100_000, read_length_min=150, read_length_max=150, coverage=30.0, paired=1,
                insert_mean=350.0, insert_sd=50.0, q_first=30.0, q_last=20.0, seed=14), {}),
    # variable lengths 100..250 with qualities (k forced to 20)
    "varlen_q": (dict(genome_length=80_000, read_length_min=100, read_length_max=250, coverage=30.0,
                      q_first=30.0, q_last=20.0, seed=15), dict(use_quality_scores=1, kmerlength=20)),
    # N bases -> ambiguous reads (edits disabled for them)
    "nbases": (dict(genome_length=60_000, read_length_min=150, read_length_max=150, coverage=30.0,
                    q_first=28.0, q_last=20.0, n_rate=0.002, seed=16), {}),
    "nbases_exclude": (dict(genome_length=60_000, read_length_min=150, read_length_max=150, coverage=30.0,
                            q_first=28.0, q_last=20.0, n_rate=0.002, seed=16), dict(exclude_ambiguous=1)),
    # tandem repeats: low-complexity reads with many equally scored alignment shifts
    "se_tandem": (dict(genome_length=60_000, tandem_fraction=0.3, tandem_period_max=6, read_length_min=150,
                       read_length_max=150, coverage=30.0, q_first=30.0, q_last=20.0, seed=18), {}),
    # 280..330 bp (16 register columns per lane in the wave MSA kernel), k = 32 as the
    # reference picks for reads of 160 bp and more
    "len300_q": (dict(genome_length=60_000, read_length_min=280, read_length_max=330, coverage=30.0,
                      q_first=30.0, q_last=20.0, seed=19), dict(use_quality_scores=1, kmerlength=32)),
    # 360..420 bp: the whole MSA is wider than the register kernel (workgroup kernel for the
    # intermediate dumps); the plain correction's anchor-column MSA has 8 register columns
    # per lane; alignment by the wavefront-per-pair kernel (planes wider than 11 words)
    "len400": (dict(genome_length=60_000, read_length_min=360, read_length_max=420, coverage=30.0,
                    q_first=30.0, q_last=20.0, seed=20), dict(kmerlength=32)),
    # 600..760 bp, qualities: anchor-column MSA with 12 register columns per lane
    "len700_q": (dict(genome_length=40_000, read_length_min=600, read_length_max=760, coverage=25.0,
                      q_first=30.0, q_last=20.0, seed=22),
                 dict(use_quality_scores=1, kmerlength=32, estimated_coverage=25.0)),
    # 850..1000 bp, no qualities: anchor-column MSA with 16 register columns per lane
    "len900": (dict(genome_length=40_000, read_length_min=850, read_length_max=1000, coverage=25.0,
                    q_first=30.0, q_last=20.0, seed=23), dict(kmerlength=32, estimated_coverage=25.0)),
    # short reads below k and near the minimum overlap
    "short": (dict(genome_length=20_000, read_length_min=15, read_length_max=60, coverage=40.0,
                   q_first=30.0, q_last=24.0, seed=17), dict(kmerlength=20, estimated_coverage=40.0)),
}


# Workloads checked at their full size (the benchmark's sizes; inputs from sim.WORKLOADS
# with the default seed): workload -> engine/oracle options. The records of the compiled
# reference are pinned per read-id slice in tests/golden/full_<workload>.json.
FULL_WORKLOADS = {
    # 1.0M x 150 bp SE, 5 Mbp with duplicated segments
    "se_5mbp_150bp_30x_dup": dict(use_quality_scores=1),
    # 2.0M pairs 2x150 bp, 20 Mbp with 20-copy repeat families
    "pe_20mbp_2x150bp_30x_rep": dict(use_quality_scores=1, paired=1),
    # 8.57M SE reads of 100..250 bp, 50 Mbp with 10-copy repeat families; k = 20
    "se_50mbp_100to250bp_30x_rep": dict(use_quality_scores=1, kmerlength=20),
}
FULL_SLICES = 16


def full_case(workload: str):
    s = sim.workload(workload)
    rs = reads.storage_from_sim(s, use_qualities=True)
    return s, rs, dict(FULL_WORKLOADS[workload])


def slice_digests(data: bytes, n: int, slices: int = FULL_SLICES):
    """Records (in read-id order, at most one per read) cut into `slices` equal read-id
    ranges of [0, n): per range [first, end, record count, SHA-256 of its record bytes]."""
    import hashlib
    bounds = [n * i // slices for i in range(slices + 1)]
    out = []
    p = 0
    last = -1
    for i in range(slices):
        lo, hi = bounds[i], bounds[i + 1]
        h = hashlib.sha256()
        count = 0
        start = p
        while p < len(data):
            rid = int.from_bytes(data[p:p + 4], "little")
            if rid >= hi:
                break
            assert rid > last, "records must be in ascending read-id order"
            last = rid
            p += 8 + (int.from_bytes(data[p + 4:p + 8], "little") & ((1 << 29) - 1))
            count += 1
        h.update(data[start:p])
        out.append([lo, hi, count, h.hexdigest()])
    assert p == len(data)
    return out


@functools.lru_cache(maxsize=None)
def case(name):
    simkw, opts = CASES[name]
    s = sim.simulate(**simkw)
    use_q = bool(opts.get("use_quality_scores", 0))
    rs = reads.storage_from_sim(s, use_qualities=use_q)
    o = dict(opts)
    o.setdefault("paired", int(rs.paired))
    o.setdefault("use_quality_scores", int(use_q))
    return s, rs, o


def parse_records(data: bytes):
    out = []
    p = 0
    while p < len(data):
        rid = int.from_bytes(data[p:p + 4], "little")
        flags = int.from_bytes(data[p + 4:p + 8], "little")
        nb = flags & ((1 << 29) - 1)
        out.append((rid, flags, data[p + 8:p + 8 + nb]))
        p += 8 + nb
    return out


def apply_records(rs, data: bytes) -> list[bytes]:
    """Corrected sequences of all reads (uncorrected where no record exists)."""
    seqs = [rs.decoded(i) for i in range(rs.n)]
    for rid, flags, payload in parse_records(data):
        use_edits = (flags >> 30) & 1
        if use_edits:
            ne = int.from_bytes(payload[0:4], "little", signed=True)
            pos = np.frombuffer(payload[4:4 + 4 * ne], np.int32)
            bases = payload[4 + 4 * ne:4 + 5 * ne]
            s = bytearray(seqs[rid])
            for p, b in zip(pos, bases):
                s[int(p)] = b
            seqs[rid] = bytes(s)
        else:
            ln = int.from_bytes(payload[0:4], "little", signed=True)
            seqs[rid] = bytes(payload[4:4 + ln])
    return seqs


def error_rate(sim_reads, seqs) -> float:
    mism = 0
    total = 0
    for i, s in enumerate(seqs):
        L = int(sim_reads.lengths[i])
        t = sim_reads.truth[i, :L].tobytes()
        a = np.frombuffer(s, np.uint8)
        b = np.frombuffer(t, np.uint8)
        mism += int((a != b).sum())
        total += L
    return mism / max(total, 1)
