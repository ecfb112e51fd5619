"""CPU restatement of the correction stage's output construction (test infrastructure only:
tests use it as the checker of the device merge; nothing in the product imports it).

What it restates, all from the reference:
- the FASTQ reader for 4-line records (kseqpp KseqPP::next, include/kseqpp/kseqpp.hpp:55-120,
  line reads :260-290): the header is the line after '@'; a '\\r' ending a header, sequence or
  quality string of more than one character is dropped; the '+' line's text is skipped;
- the ordering of records by read id (sortSerializedResultsByReadIdAscending,
  include/sortserializedresults.hpp:17-78);
- record decoding (TempCorrectedSequence::decode, include/correctedsequence.hpp:437-484);
- the combine rule (combineMultipleCorrectionResults1_rawtcs2,
  src/correctionresultoutput.cpp:1256-1550) and the optional " care:q=N" header labels
  (:929-943), which only reads up to the highest read id with a record receive;
- the queue writer's record text "@" header "\\n" sequence "\\n+\\n" quality "\\n"
  (src/correctionresultoutput.cpp:726-760), reads of the second input file of a pair going to
  the second queue (:727-729 with src/peqc.cpp:135-148);
- with no records at all the merge hands over nothing (:466-500) and QC reads the input file
  itself (src/seqc.cpp:662-672, src/peqc.cpp:1395-1405): the stream QC sees is the input text.

Pinned against the reference's own merge compiled from source (oracle/_ref) in
tests/test_output_oracle.py.
"""
from __future__ import annotations

import numpy as np

ANCHOR, CANDIDATE = 0, 1


def _strip_cr(line: bytes) -> bytes:
    return line[:-1] if len(line) > 1 and line.endswith(b"\r") else line


def parse_fastq(text: bytes):
    """[(header, sequence, quality)] of 4-line FASTQ text."""
    lines = text.split(b"\n")
    while lines and lines[-1] in (b"", b"\r"):
        lines.pop()
    if len(lines) % 4:
        raise ValueError("not 4-line FASTQ")
    out = []
    for i in range(0, len(lines), 4):
        h, s, p, q = lines[i:i + 4]
        if not h.startswith(b"@") or not p.startswith(b"+"):
            raise ValueError("malformed record")
        s, q = _strip_cr(s), _strip_cr(q)
        if len(s) != len(q):
            raise ValueError("quality length")
        out.append((_strip_cr(h[1:]), s, q))
    return out


def decode_records(data: bytes):
    """Records in stream order: dicts with read_id, hq, use_edits, type, edits [(pos, base)],
    sequence (bytes or None), shift."""
    out = []
    p = 0
    while p < len(data):
        rid = int.from_bytes(data[p:p + 4], "little")
        flags = int.from_bytes(data[p + 4:p + 8], "little")
        nb = flags & ((1 << 29) - 1)
        body = data[p + 8:p + 8 + nb]
        r = dict(read_id=rid, hq=bool(flags >> 31 & 1), use_edits=bool(flags >> 30 & 1), type=flags >> 29 & 1,
                 edits=[], sequence=None, shift=0)
        q = 0
        if r["use_edits"]:
            ne = int.from_bytes(body[0:4], "little", signed=True)
            pos = np.frombuffer(body[4:4 + 4 * ne], "<i4")
            bases = body[4 + 4 * ne:4 + 5 * ne]
            r["edits"] = [(int(a), bases[i:i + 1]) for i, a in enumerate(pos)]
            q = 4 + 5 * ne
        else:
            ln = int.from_bytes(body[0:4], "little", signed=True)
            r["sequence"] = bytes(body[4:4 + ln])
            q = 4 + ln
        if r["type"] != ANCHOR:
            r["shift"] = int.from_bytes(body[q:q + 4], "little", signed=True)
        out.append(r)
        p += 8 + nb
    return out


def _apply(seq: bytearray, r):
    if r["use_edits"]:
        for pos, b in r["edits"]:
            seq[pos:pos + 1] = b
        return seq
    return bytearray(r["sequence"])


def _equal_one_has_edits(orig: bytes, e, s) -> bool:
    seq2 = s["sequence"]
    edits = e["edits"]
    k = 0
    pos = 0
    n = len(orig)
    while pos != n and k != len(edits):
        if pos != edits[k][0]:
            if orig[pos] != seq2[pos]:
                return False
        else:
            if edits[k][1][0] != seq2[pos]:
                return False
            k += 1
        pos += 1
    while pos != n:
        if orig[pos] != seq2[pos]:
            return False
        pos += 1
    return True


def _equal(orig: bytes, a, b) -> bool:
    la = len(orig) if a["use_edits"] else len(a["sequence"])
    lb = len(orig) if b["use_edits"] else len(b["sequence"])
    if la != lb:
        return False
    if a["use_edits"] and b["use_edits"]:
        return a["edits"] == b["edits"]
    if not a["use_edits"] and not b["use_edits"]:
        return a["sequence"] == b["sequence"]
    e, s = (a, b) if a["use_edits"] else (b, a)
    if not e["edits"]:
        return orig == s["sequence"]
    return _equal_one_has_edits(orig, e, s)


def combine(results, orig: bytes):
    """(sequence, label) of one read: label 0 = not corrected, 1 = anchor alone (LQ),
    2 = LQ with agreeing candidates' results, 3 = HQ anchor."""
    if not results:
        return orig, 0
    anchor = next((r for r in results if r["type"] == ANCHOR), None)
    if anchor is None:
        return orig, 0
    if anchor["hq"]:
        return bytes(_apply(bytearray(orig), anchor)), 3
    if len(results) >= 3:
        if all(_equal(orig, results[0], r) for r in results[1:]):
            return bytes(_apply(bytearray(orig), results[0])), 2
        return orig, 0
    return bytes(_apply(bytearray(orig), anchor)), 1


def merge(texts, records: bytes, paired: bool, labels: bool = False):
    """Output text per input file (what the reference's merge hands to QC)."""
    recs = decode_records(records)
    if not recs:
        return [bytes(t) for t in texts]
    files = [parse_fastq(t) for t in texts]
    if paired:
        reads = [None] * (2 * len(files[0]))
        reads[0::2] = [(0, x) for x in files[0]]
        reads[1::2] = [(1, x) for x in files[1]]
    else:
        reads = [(0, x) for x in files[0]]
    by_id = {}
    for r in sorted(recs, key=lambda r: r["read_id"]):     # stable: stream order within a read
        by_id.setdefault(r["read_id"], []).append(r)
    outs = [bytearray() for _ in texts]
    tag = {0: b" care:q=0", 1: b" care:q=1", 2: b" care:q=2", 3: b" care:q=3"}
    last = max(by_id)
    for rid, (f, (h, s, q)) in enumerate(reads):
        seq, lab = combine(by_id.get(rid, []), s)
        # the merge stops visiting reads once the last record is used (:901-906): reads
        # after the highest read id with a record keep their header unlabelled
        if labels and rid <= last:
            h = h + tag[lab]
        outs[f] += b"@" + h + b"\n" + seq + b"\n+\n" + q + b"\n"
    return [bytes(o) for o in outs]
