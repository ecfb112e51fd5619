"""Python view of the seeded read simulator (``include/care_sim.h``).

The workload presets below are the correction workloads this engine is measured on.
Each preset is a pure function of its seed, so inputs are regenerated where they are
needed instead of being stored.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import build as _build

_LIB = None


class SimConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("genome_length", ctypes.c_uint64),
        ("gc_content", ctypes.c_double),
        ("dup_fraction", ctypes.c_double),
        ("dup_segment_length", ctypes.c_int),
        ("dup_copies", ctypes.c_int),
        ("dup_divergence", ctypes.c_double),
        ("repeat_fraction", ctypes.c_double),
        ("repeat_unit_length", ctypes.c_int),
        ("repeat_copies", ctypes.c_int),
        ("repeat_divergence", ctypes.c_double),
        ("read_length_min", ctypes.c_int),
        ("read_length_max", ctypes.c_int),
        ("coverage", ctypes.c_double),
        ("num_reads_override", ctypes.c_uint64),
        ("paired", ctypes.c_int),
        ("insert_mean", ctypes.c_double),
        ("insert_sd", ctypes.c_double),
        ("q_first", ctypes.c_double),
        ("q_last", ctypes.c_double),
        ("q_sd", ctypes.c_double),
        ("q_min", ctypes.c_int),
        ("q_max", ctypes.c_int),
        ("n_rate", ctypes.c_double),
        ("tandem_fraction", ctypes.c_double),
        ("tandem_period_max", ctypes.c_int),
        ("per_read_streams", ctypes.c_int),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_build.LIBDIR, "libcaresim.so")
        if not os.path.exists(path):
            _build.build_sim()
        lib = ctypes.CDLL(path)
        lib.care_sim_default_config.argtypes = [ctypes.POINTER(SimConfig)]
        lib.care_sim_run.argtypes = [ctypes.POINTER(SimConfig), ctypes.POINTER(ctypes.c_void_p)]
        lib.care_sim_run.restype = ctypes.c_int
        lib.care_sim_num_reads.argtypes = [ctypes.c_void_p]
        lib.care_sim_num_reads.restype = ctypes.c_uint64
        lib.care_sim_max_length.argtypes = [ctypes.c_void_p]
        lib.care_sim_max_length.restype = ctypes.c_int
        lib.care_sim_copy.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]
        lib.care_sim_copy.restype = ctypes.c_int
        lib.care_sim_free.argtypes = [ctypes.c_void_p]
        lib.care_sim_fastq.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.c_size_t]
        lib.care_sim_fastq.restype = ctypes.c_size_t
        _LIB = lib
    return _LIB


@dataclass
class SimReads:
    """Simulated reads as row-major byte matrices (``pitch`` bytes per read)."""
    seqs: np.ndarray        # uint8 [n, pitch] ASCII bases, zero padded
    quals: np.ndarray       # uint8 [n, pitch] Phred+33, zero padded
    truth: np.ndarray       # uint8 [n, pitch] error-free read
    lengths: np.ndarray     # int32 [n]
    positions: np.ndarray   # uint64 [n]
    strands: np.ndarray     # int32 [n]
    paired: bool = False
    config: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.lengths.shape[0])

    @property
    def max_length(self) -> int:
        return int(self.lengths.max()) if self.n else 0


def simulate(**kw) -> SimReads:
    lib = _lib()
    cfg = SimConfig()
    lib.care_sim_default_config(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown simulator option {k}")
        setattr(cfg, k, v)
    handle = ctypes.c_void_p()
    rc = lib.care_sim_run(ctypes.byref(cfg), ctypes.byref(handle))
    if rc != 0:
        raise RuntimeError(f"care_sim_run failed with {rc}")
    try:
        n = lib.care_sim_num_reads(handle)
        maxlen = lib.care_sim_max_length(handle)
        seqs = np.empty((n, maxlen), np.uint8)        # every byte is written by care_sim_copy
        quals = np.empty((n, maxlen), np.uint8)
        truth = np.empty((n, maxlen), np.uint8)
        lengths = np.zeros(n, np.int32)
        positions = np.zeros(n, np.uint64)
        strands = np.zeros(n, np.int32)
        rc = lib.care_sim_copy(handle, 0, n, maxlen, seqs.ctypes.data, quals.ctypes.data, truth.ctypes.data,
                               lengths.ctypes.data, positions.ctypes.data, strands.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"care_sim_copy failed with {rc}")
    finally:
        lib.care_sim_free(handle)
    return SimReads(seqs, quals, truth, lengths, positions, strands, bool(cfg.paired), dict(kw))


# Correction workloads (sizes as named). Qualities fall linearly from the first to
# the last base; substitutions occur with probability 10^(-Q/10). The three large
# workloads draw each read from its own generator stream (simulated on all host threads).
WORKLOADS = {
    # plain SE: 200 kbp uniform genome, 150 bp reads at 30x (~40k reads)
    "se_200kbp_150bp_30x": dict(genome_length=200_000, read_length_min=150, read_length_max=150,
                                coverage=30.0, q_first=36.0, q_last=28.0),
    # batched SE: 5 Mbp, 50% GC, 2% of it in 1 kbp segments copied twice more at 0.5%
    # divergence; 150 bp reads at 30x (~1.0M reads)
    "se_5mbp_150bp_30x_dup": dict(genome_length=5_000_000, gc_content=0.5, dup_fraction=0.02,
                                  dup_segment_length=1000, dup_copies=2, dup_divergence=0.005,
                                  read_length_min=150, read_length_max=150, coverage=30.0,
                                  q_first=36.0, q_last=28.0, per_read_streams=1),
    # batched PE: 20 Mbp, 10% in 20-copy 2 kbp repeat families at 1% divergence;
    # 2x150 bp pairs, insert N(350,50), 30x; steeper 3' quality drop (Q24 at the end)
    "pe_20mbp_2x150bp_30x_rep": dict(genome_length=20_000_000, repeat_fraction=0.10,
                                     repeat_unit_length=2000, repeat_copies=20, repeat_divergence=0.01,
                                     read_length_min=150, read_length_max=150, coverage=30.0, paired=1,
                                     insert_mean=350.0, insert_sd=50.0, q_first=36.0, q_last=24.0,
                                     per_read_streams=1),
    # learned inference at scale: 50 Mbp, 8% in 10-copy repeats; lengths U[100,250] at 30x
    "se_50mbp_100to250bp_30x_rep": dict(genome_length=50_000_000, repeat_fraction=0.08,
                                        repeat_unit_length=2000, repeat_copies=10, repeat_divergence=0.01,
                                        read_length_min=100, read_length_max=250, coverage=30.0,
                                        q_first=36.0, q_last=28.0, per_read_streams=1),
}


def workload(name: str, seed: int = 12345, **overrides) -> SimReads:
    kw = dict(WORKLOADS[name])
    kw.update(overrides)
    kw["seed"] = seed
    return simulate(**kw)


def fastq_texts(s: SimReads) -> list[np.ndarray]:
    """4-line FASTQ text of simulated reads (uint8 arrays): one text, or R1/R2 texts for
    pairs (reads 2i and 2i+1 are mates). Read i is written as ``@r<i>``."""
    lib = _lib()
    n = int(s.lengths.shape[0])
    seqs = np.ascontiguousarray(s.seqs)
    quals = np.ascontiguousarray(s.quals)
    lengths = np.ascontiguousarray(s.lengths, np.int32)
    parts = [(0, 1, n)] if not s.paired else [(0, 2, n // 2), (1, 2, n // 2)]
    out = []
    for first, step, count in parts:
        size = lib.care_sim_fastq(seqs.ctypes.data, quals.ctypes.data, lengths.ctypes.data, seqs.shape[1],
                                  first, step, count, None, 0)
        buf = np.empty(max(1, size), np.uint8)
        got = lib.care_sim_fastq(seqs.ctypes.data, quals.ctypes.data, lengths.ctypes.data, seqs.shape[1],
                                 first, step, count, buf.ctypes.data, size)
        if got != size:
            raise RuntimeError("care_sim_fastq failed")
        out.append(buf[:size])
    return out
