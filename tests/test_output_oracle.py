"""The output-merge restatement (oracle/output_merge.py) against the reference's own merge
compiled from source (sort by read id, combine rule, queue writer; CPU only).

Cases: the golden records of simulated read sets (single and paired end), FASTQ texts
with CRLF line ends, '+' lines with text, lower-case and N bases; hand-made record sets
that reach every branch of the combine rule (HQ anchor, anchor alone, three or more
agreeing or disagreeing results, candidate results only, full sequences against edits,
unequal lengths), given out of read-id order, with and without the " care:q=N" labels; and
no records at all.
"""
import os
import random
import struct

import numpy as np
import pytest

import oracle
from oracle import output_merge
from tests import fastq_data, helpers

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def encode_record(read_id, hq=False, use_edits=True, type_=0, edits=(), sequence=b"", shift=0):
    """One serialized record (EncodedTempCorrectedSequence, include/correctedsequence.hpp:299-380)."""
    if use_edits:
        body = struct.pack("<i", len(edits)) + b"".join(struct.pack("<i", p) for p, _ in edits)
        body += b"".join(b for _, b in edits)
    else:
        body = struct.pack("<i", len(sequence)) + sequence
    if type_:
        body += struct.pack("<i", shift)
    flags = (int(hq) << 31) | (int(use_edits) << 30) | (int(type_) << 29) | len(body)
    return struct.pack("<II", read_id, flags) + body


def sim_texts(name, mangle=None):
    s, rs, o = helpers.case(name)
    data = dict(seq_matrix=s.seqs, qual_matrix=s.quals, lengths=s.lengths)
    n = rs.n
    parts = [range(n)] if not rs.paired else [range(0, n, 2), range(1, n, 2)]
    texts = [fastq_data.text(data, p) for p in parts]
    if mangle:
        texts = [mangle(t) for t in texts]
    return texts, rs.paired


def run_both(tmp_path, texts, records, paired, labels):
    paths = []
    for i, t in enumerate(texts):
        p = str(tmp_path / f"in{i}.fq")
        open(p, "wb").write(t)
        paths.append(p)
    ref, changed = oracle.ref_output_fastq(paths[0], paths[1] if paired else None, records, labels=labels,
                                           tmpdir=str(tmp_path))
    ours = output_merge.merge(texts, records, paired, labels)
    return ref, changed, ours


def _mangle_crlf(t):
    return t.replace(b"\n+\n", b"\n+ and some text\n").replace(b"\n", b"\r\n")


def _mangle_case(t):
    lines = t.split(b"\n")
    for i in range(1, len(lines), 40):           # every tenth record's sequence in lower case
        lines[i] = lines[i].lower()
    for i in range(5, len(lines), 28):           # an N in every seventh record
        if lines[i]:
            lines[i] = b"N" + lines[i][1:]
    return b"\n".join(lines)


@pytest.mark.parametrize("name", ["se_q", "pe", "varlen_q"])
@pytest.mark.parametrize("variant", ["plain", "crlf", "case"])
@pytest.mark.parametrize("labels", [False, True])
def test_merge_of_golden_records_matches_reference(tmp_path, name, variant, labels):
    mangle = {"plain": None, "crlf": _mangle_crlf, "case": _mangle_case}[variant]
    texts, paired = sim_texts(name, mangle)
    records = np.load(os.path.join(GOLDEN, f"records_{name}.npz"))["records"].tobytes()
    ref, changed, ours = run_both(tmp_path, texts, records, paired, labels)
    assert changed == len(output_merge.decode_records(records)) > 0
    assert ref == ours


def combine_rule_case(n_reads=40, seed=0):
    """FASTQ text of n_reads reads and a record set reaching every branch of the combine rule."""
    data = fastq_data.make(seed=seed, n_reads=n_reads, paired=False, min_len=40, max_len=60)
    seqs = [data["seq_matrix"][i, :data["lengths"][i]].tobytes() for i in range(n_reads)]
    rng = random.Random(seed)

    def edits_for(i, k, salt=0):
        r = random.Random(1000 * i + salt)
        pos = sorted(r.sample(range(len(seqs[i])), k))
        return [(p, bytes([b"ACGT"[(b"ACGT".find(seqs[i][p:p + 1].upper()) + 1) % 4]])) for p in pos]

    def applied(i, edits):
        s = bytearray(seqs[i])
        for p, b in edits:
            s[p:p + 1] = b
        return bytes(s)

    per_read = {}
    e = edits_for(0, 3); per_read[0] = [encode_record(0, hq=True, edits=e)]
    e = edits_for(1, 2); per_read[1] = [encode_record(1, edits=e)]
    e = edits_for(2, 2)
    per_read[2] = [encode_record(2, edits=e), encode_record(2, type_=1, edits=e, shift=3),
                   encode_record(2, type_=1, use_edits=False, sequence=applied(2, e), shift=-2)]
    e = edits_for(3, 2)
    per_read[3] = [encode_record(3, edits=e), encode_record(3, type_=1, edits=e, shift=1),
                   encode_record(3, type_=1, edits=edits_for(3, 2, salt=1), shift=1)]
    per_read[4] = [encode_record(4, type_=1, edits=edits_for(4, 1), shift=5)] * 3
    per_read[5] = [encode_record(5, hq=True, use_edits=False, sequence=applied(5, edits_for(5, 4)))]
    e = edits_for(6, 1)
    per_read[6] = [encode_record(6, type_=1, use_edits=False, sequence=applied(6, e), shift=2),
                   encode_record(6, edits=e), encode_record(6, type_=1, edits=e, shift=4),
                   encode_record(6, type_=1, use_edits=False, sequence=applied(6, e), shift=7)]
    per_read[7] = [encode_record(7, edits=[]), encode_record(7, type_=1, use_edits=False, sequence=seqs[7]),
                   encode_record(7, type_=1, edits=[])]
    per_read[8] = [encode_record(8, edits=edits_for(8, 1)),
                   encode_record(8, type_=1, use_edits=False, sequence=seqs[8][:-1]),
                   encode_record(8, type_=1, edits=edits_for(8, 1))]
    per_read[9] = [encode_record(9, edits=edits_for(9, 2)), encode_record(9, type_=1, edits=edits_for(9, 2))]
    per_read[10] = [encode_record(10, use_edits=False, sequence=seqs[10][:-3] + b"ACG")]
    for i in range(11, n_reads - 8, 3):            # a spread of plain anchor records
        per_read[i] = [encode_record(i, hq=bool(i % 2), edits=edits_for(i, 1 + i % 3))]
    # records handed over out of read-id order (stream order kept within a read)
    order = list(per_read)
    rng.shuffle(order)
    records = b"".join(b"".join(per_read[i]) for i in order)
    return data, records


@pytest.mark.parametrize("labels", [False, True])
@pytest.mark.parametrize("variant", ["plain", "crlf", "case"])
def test_combine_rule_branches_match_reference(tmp_path, labels, variant):
    data, records = combine_rule_case()
    mangle = {"plain": None, "crlf": _mangle_crlf, "case": _mangle_case}[variant]
    text = fastq_data.text(data, range(len(data["lengths"])))
    texts = [mangle(text) if mangle else text]
    ref, changed, ours = run_both(tmp_path, texts, records, False, labels)
    assert ref == ours
    if labels:
        out = ours[0].split(b"\n")
        heads = [out[i] for i in range(0, len(out) - 1, 4)]
        got = {int(h.split(b" ")[0][2:]): h.split(b"care:q=")[1] if b"care:q=" in h else None for h in heads}
        assert (got[0], got[1], got[2], got[3], got[4], got[5]) == (b"3", b"1", b"2", b"0", b"0", b"3")
        assert got[len(heads) - 1] is None          # past the last record: no label


def test_paired_combine_rule_matches_reference(tmp_path):
    data, records = combine_rule_case(n_reads=60, seed=3)
    texts = [fastq_data.text(data, range(0, 60, 2)), fastq_data.text(data, range(1, 60, 2))]
    for labels in (False, True):
        ref, changed, ours = run_both(tmp_path, texts, records, True, labels)
        assert ref == ours


def test_no_records_hands_over_nothing(tmp_path):
    texts, paired = sim_texts("se_q", _mangle_crlf)
    ref, changed, ours = run_both(tmp_path, texts, b"", paired, False)
    assert changed == 0 and ref == [b""]
    assert ours == texts        # QC then reads the input text itself
