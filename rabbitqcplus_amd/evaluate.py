"""Post-correction base error rates against the simulator's truth.

Applies serialized correction records (read id u32, flags u32 with bit 30 = edits,
low 29 bits = payload bytes; payload = edit count + positions + bases, or length + full
sequence: include/correctedsequence.hpp:299-387) to the decoded reads and counts
mismatches against the true read sequences.
"""
from __future__ import annotations

import numpy as np


def decode_all(rs) -> np.ndarray:
    """ASCII bases [n][16 * pitch_ints] of a ReadStorage (bytes past a read's length are 'A')."""
    shifts = (30 - 2 * np.arange(16, dtype=np.uint32)).astype(np.uint32)
    codes = ((rs.enc[:, :, None] >> shifts[None, None, :]) & 3).reshape(rs.n, -1).astype(np.uint8)
    return np.frombuffer(b"ACGT", np.uint8)[codes]


def apply_records(seqs: np.ndarray, data: bytes) -> np.ndarray:
    """Corrected copy of `seqs` ([n][w] ASCII) after the records in `data`."""
    out = seqs.copy()
    p = 0
    mv = memoryview(data)
    while p < len(data):
        rid = int.from_bytes(mv[p:p + 4], "little")
        flags = int.from_bytes(mv[p + 4:p + 8], "little")
        nb = flags & ((1 << 29) - 1)
        pl = mv[p + 8:p + 8 + nb]
        if (flags >> 30) & 1:
            ne = int.from_bytes(pl[0:4], "little")
            pos = np.frombuffer(pl[4:4 + 4 * ne], np.int32)
            out[rid, pos] = np.frombuffer(pl[4 + 4 * ne:4 + 5 * ne], np.uint8)
        else:
            ln = int.from_bytes(pl[0:4], "little")
            out[rid, :ln] = np.frombuffer(pl[4:4 + ln], np.uint8)
        p += 8 + nb
    return out


def error_counts(s, rs, data: bytes) -> dict:
    """bases, errors before/after, miscorrected (right before, wrong after) and fixed
    (wrong before, right after) base counts."""
    seqs = decode_all(rs)
    w = min(seqs.shape[1], s.truth.shape[1])
    valid = np.arange(w)[None, :] < rs.lengths[:, None]
    truth = s.truth[:, :w]
    before = (seqs[:, :w] != truth) & valid
    after = (apply_records(seqs, data)[:, :w] != truth) & valid
    return dict(bases=int(rs.lengths.sum()), before=int(before.sum()), after=int(after.sum()),
                miscorrected=int((after & ~before).sum()), fixed=int((before & ~after).sum()))


def error_rates(s, rs, data: bytes):
    c = error_counts(s, rs, data)
    return c["before"] / c["bases"], c["after"] / c["bases"]
