"""Seeded FASTQ test data: reads with N and other non-ACGT characters, lower-case bases,
variable lengths, qualities; written as 4-line FASTQ (one file, or R1/R2 for pairs)."""
from __future__ import annotations

import os

import numpy as np


def make(seed: int, n_reads: int, paired: bool, min_len: int = 60, max_len: int = 150):
    rng = np.random.default_rng(seed)
    lens = rng.integers(min_len, max_len + 1, n_reads).astype(np.int32)
    alphabet = np.frombuffer(b"ACGT", np.uint8)
    seq = np.zeros((n_reads, max_len), np.uint8)
    qual = np.zeros((n_reads, max_len), np.uint8)
    for i in range(n_reads):
        L = int(lens[i])
        s = alphabet[rng.integers(0, 4, L)]
        if rng.random() < 0.05:                       # some lower-case reads
            s = np.where(rng.random(L) < 0.5, s + 32, s).astype(np.uint8)
        if rng.random() < 0.02:                       # ambiguous reads: N and IUPAC codes
            k = int(rng.integers(1, 4))
            pos = rng.choice(L, k, replace=False)
            s[pos] = np.frombuffer(b"NRYn", np.uint8)[rng.integers(0, 4, k)]
        seq[i, :L] = s
        qual[i, :L] = rng.integers(35, 74, L)
    return dict(seq_matrix=seq, qual_matrix=qual, lengths=lens)


def records(data, idx):
    seqs = [data["seq_matrix"][i, :data["lengths"][i]].tobytes() for i in idx]
    quals = [data["qual_matrix"][i, :data["lengths"][i]].tobytes() for i in idx]
    return seqs, quals


def text(data, idx) -> bytes:
    seqs, quals = records(data, idx)
    return b"".join(b"@r%d\n" % i + s + b"\n+\n" + q + b"\n" for i, s, q in zip(idx, seqs, quals))


def write(tmp_path, data, paired):
    n = len(data["lengths"])
    if not paired:
        p = os.path.join(str(tmp_path), "reads.fq")
        open(p, "wb").write(text(data, range(n)))
        return (p,)
    p1 = os.path.join(str(tmp_path), "r1.fq")
    p2 = os.path.join(str(tmp_path), "r2.fq")
    open(p1, "wb").write(text(data, range(0, n, 2)))
    open(p2, "wb").write(text(data, range(1, n, 2)))
    return (p1, p2)


def parse_reads_file(path):
    """Fields of a preprocessed-reads file (ChunkedReadStorage::saveToFile,
    include/chunkedreadstorage.hpp:246-398; lengths include/lengthstorage.hpp:184-204)."""
    import struct
    b = open(path, "rb").read()
    p = 0
    n, upper, lower = struct.unpack_from("<Qii", b, p); p += 16
    has_q = b[p]; p += 1
    (qbits,) = struct.unpack_from("<i", b, p); p += 4
    sections = struct.unpack_from("<QQQQ", b, p); p += 32
    dtb, bits, mn, mx, mask, count, lwords, lbytes = struct.unpack_from("<iiiiIqQQ", b, p); p += 44
    lw = np.frombuffer(b, "<u4", lwords, p).copy(); p += lbytes
    sp, sn = struct.unpack_from("<QQ", b, p); p += 16
    seq = np.frombuffer(b, "<u4", sn, p).copy(); p += 4 * sn
    qp, qn = struct.unpack_from("<QQ", b, p); p += 16
    qual = np.frombuffer(b, "<u4", qn, p).copy(); p += 4 * qn
    (na,) = struct.unpack_from("<Q", b, p); p += 8
    amb = np.frombuffer(b, "<u4", na, p).copy(); p += 4 * na
    assert p == len(b)
    if bits:
        pos = np.arange(n, dtype=np.int64)[:, None] * bits + np.arange(bits)[None, :]
        bit = (lw[pos >> 5] >> (31 - (pos & 31)).astype(np.uint32)) & 1
        lengths = mn + (bit.astype(np.int64) << np.arange(bits - 1, -1, -1)).sum(1)
    else:
        lengths = np.full(n, mn, np.int64)
    return dict(n=n, upper=upper, lower=lower, has_q=has_q, qbits=qbits, sections=sections, dtb=dtb, bits=bits,
                min=mn, max=mx, mask=mask, count=count, length_words=lw, lengths=lengths.astype(np.int32),
                seq_pitch=sp, enc=seq.reshape(n, sp) if sp else seq, qual_pitch=qp,
                quals=qual.view(np.uint8).reshape(n, 4 * qp) if qp else None, ambiguous_ids=amb)
