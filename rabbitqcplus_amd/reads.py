"""Host-side read storage: the input contract of the correction engine.

Reads enter the engine as 2-bit rows (A=0 C=1 G=2 T=3, 16 bases per u32, first base
in the two most significant bits, last word left-aligned; reference
include/sequencehelpers.hpp:154-187), a length per read, optional Phred+33
qualities in rows of ``max_length`` bytes, and an ambiguity flag per read.

Bases other than ACGT (case-insensitive) are replaced cyclically by A,C,G,T, the
counter restarting for every 65,536-read parser batch, and the read is flagged
ambiguous (include/chunkedreadstorageconstruction.hpp:67-92, batch size :106).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

PARSER_BATCH = 65536
_CODE = np.full(256, 255, np.uint8)
for _c, _v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
    _CODE[_c] = _v
_DECODE = np.frombuffer(b"ACGT", np.uint8)


def ints_per_row(max_length: int) -> int:
    """u32 words of one 2-bit row (getEncodedNumInts2Bit)."""
    return (int(max_length) + 15) // 16


@dataclass
class ReadStorage:
    enc: np.ndarray          # uint32 [n, pitch_ints]
    lengths: np.ndarray      # int32 [n]
    quals: np.ndarray | None # uint8 [n, max_length] or None
    ambiguous: np.ndarray    # uint8 [n]
    paired: bool = False

    @property
    def n(self) -> int:
        return int(self.lengths.shape[0])

    @property
    def pitch_ints(self) -> int:
        return int(self.enc.shape[1])

    @property
    def max_length(self) -> int:
        return int(self.lengths.max()) if self.n else 0

    def decoded(self, i: int) -> bytes:
        return decode_row(self.enc[i], int(self.lengths[i]))


def replace_ambiguous(seqs: np.ndarray, lengths: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Returns (codes uint8 [n, pitch] in 0..3, ambiguous uint8 [n])."""
    n, pitch = seqs.shape
    codes = _CODE[seqs]
    valid_pos = np.arange(pitch)[None, :] < lengths[:, None]
    bad = (codes == 255) & valid_pos
    ambiguous = bad.any(axis=1).astype(np.uint8)
    if ambiguous.any():
        rows, cols = np.nonzero(bad)            # row-major order = file order
        batch = rows // PARSER_BATCH
        # position of each bad base within its parser batch, in file order
        first_in_batch = np.searchsorted(batch, batch, side="left")
        k = np.arange(rows.shape[0]) - first_in_batch
        codes[rows, cols] = (k % 4).astype(np.uint8)
    codes[~valid_pos] = 0
    return codes, ambiguous


def encode_codes(codes: np.ndarray, lengths: np.ndarray, words: int | None = None) -> np.ndarray:
    """2-bit rows of base codes (0..3, zero past each read): four codes per byte, first
    code in the top bits, four bytes per u32 word read most significant byte first."""
    n, pitch = codes.shape
    if words is None:
        words = ints_per_row(int(lengths.max()) if n else 0)
    padded = np.zeros((n, words * 16), np.uint8)
    w = min(pitch, words * 16)
    padded[:, :w] = codes[:, :w]
    q = padded.reshape(n, words * 4, 4)
    packed = (q[:, :, 0] << 6) | (q[:, :, 1] << 4) | (q[:, :, 2] << 2) | q[:, :, 3]
    return np.ascontiguousarray(packed).view(">u4").astype(np.uint32).reshape(n, words)


def make_storage(seqs: np.ndarray, lengths: np.ndarray, quals: np.ndarray | None = None,
                 paired: bool = False, threads: int = 0) -> ReadStorage:
    """Read storage from ASCII rows, encoded on all host threads by the library's
    ``care_hip_encode_reads`` (the rule of ``replace_ambiguous`` + ``encode_codes``)."""
    import ctypes
    from . import engine
    lib = engine.load_library()
    lengths = np.ascontiguousarray(lengths, np.int32)
    seqs = np.ascontiguousarray(seqs, np.uint8)
    n = lengths.shape[0]
    maxlen = int(lengths.max()) if n else 0
    words = ints_per_row(maxlen)
    enc = np.empty((n, max(words, 1)), np.uint32)
    ambiguous = np.empty(n, np.uint8)
    q = None
    qin = None
    if quals is not None:
        qin = np.ascontiguousarray(quals, np.uint8)
        q = np.empty((n, maxlen), np.uint8)
    rc = lib.care_hip_encode_reads(
        seqs.ctypes.data, seqs.shape[1] if seqs.ndim == 2 else 0, lengths.ctypes.data, n,
        enc.ctypes.data, enc.shape[1], ambiguous.ctypes.data,
        None if qin is None else qin.ctypes.data, 0 if qin is None else qin.shape[1],
        None if q is None else q.ctypes.data, 0 if q is None else maxlen, ctypes.c_int(threads))
    if rc != 0:
        raise ValueError(f"care_hip_encode_reads failed ({rc}): lengths exceed the given rows")
    return ReadStorage(enc[:, :words] if words else enc[:, :0], lengths, q, ambiguous, paired)


def decode_row(row: np.ndarray, length: int) -> bytes:
    words = np.asarray(row, np.uint32)
    shifts = (30 - 2 * np.arange(16, dtype=np.uint32)).astype(np.uint32)
    codes = ((words[:, None] >> shifts[None, :]) & 3).reshape(-1)[:length]
    return _DECODE[codes].tobytes()


def storage_from_sim(sim, use_qualities: bool = True) -> ReadStorage:
    return make_storage(sim.seqs, sim.lengths, sim.quals if use_qualities else None, sim.paired)
