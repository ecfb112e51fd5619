"""Output merge on the GPU (care_hip_write_fastq).

Bar: byte-exact against oracle/output_merge.py, which the CPU suite pins to the reference's
own merge compiled from source (tests/test_output_oracle.py) on the same cases: golden
records of simulated read sets with CRLF, '+' text, lower-case and N variants of their
FASTQ, with and without quality labels; a record set reaching every branch of the
combine rule, out of read-id order, single and paired end; no records; and a whole
correction run from FASTQ text to corrected FASTQ text.
"""
import os
import struct

import numpy as np
import pytest

from oracle import output_merge
from rabbitqcplus_amd import engine
from tests import fastq_data, helpers
from tests.test_output_oracle import _mangle_case, _mangle_crlf, combine_rule_case, encode_record, sim_texts

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MANGLE = {"plain": None, "crlf": _mangle_crlf, "case": _mangle_case}


def _gpu_merge(texts, records, paired, labels, **opts):
    eng = engine.CareEngine(0, paired=int(paired), use_quality_scores=1, **opts)
    try:
        eng.set_reads_fastq(*texts)
        return [eng.write_fastq(records, f, labels) for f in range(len(texts))]
    finally:
        eng.close()


@pytest.mark.parametrize("name", ["se_q", "pe", "varlen_q"])
@pytest.mark.parametrize("variant", ["plain", "crlf", "case"])
@pytest.mark.parametrize("labels", [False, True])
def test_merge_golden_records(name, variant, labels):
    texts, paired = sim_texts(name, MANGLE[variant])
    records = np.load(os.path.join(GOLDEN, f"records_{name}.npz"))["records"].tobytes()
    print(f"[merge {name} {variant} labels={labels}]", flush=True)
    assert _gpu_merge(texts, records, paired, labels) == output_merge.merge(texts, records, paired, labels)


@pytest.mark.parametrize("labels", [False, True])
@pytest.mark.parametrize("variant", ["plain", "crlf", "case"])
def test_merge_combine_rule_branches(labels, variant):
    data, records = combine_rule_case()
    text = fastq_data.text(data, range(len(data["lengths"])))
    texts = [MANGLE[variant](text) if MANGLE[variant] else text]
    assert _gpu_merge(texts, records, False, labels) == output_merge.merge(texts, records, False, labels)


def test_merge_combine_rule_paired():
    data, records = combine_rule_case(n_reads=60, seed=3)
    texts = [fastq_data.text(data, range(0, 60, 2)), fastq_data.text(data, range(1, 60, 2))]
    for labels in (False, True):
        assert _gpu_merge(texts, records, True, labels) == output_merge.merge(texts, records, True, labels)


def test_merge_without_records_returns_input():
    texts, paired = sim_texts("se_q", _mangle_crlf)
    texts = [texts[0] + b"\r\n\n"]
    assert _gpu_merge(texts, b"", paired, True) == texts


def test_correction_from_fastq_to_fastq():
    """FASTQ text -> device storage -> tables -> records -> corrected FASTQ text."""
    # lower-case bases do not change the storage, so the records are the golden ones
    texts, paired = sim_texts("pe", lambda t: b"\n".join(x.lower() if i % 40 == 1 else x
                                                         for i, x in enumerate(t.split(b"\n"))))
    golden = np.load(os.path.join(GOLDEN, "records_pe.npz"))["records"].tobytes()
    eng = engine.CareEngine(0, **helpers.case("pe")[2])
    try:
        eng.set_reads_fastq(*texts)
        eng.build_tables()
        n, records = eng.correct()
        outs = [eng.write_fastq(records, f) for f in (0, 1)]
    finally:
        eng.close()
    assert records == golden
    assert outs == output_merge.merge(texts, records, True)


def test_merge_many_reads_many_records():
    """200,000 reads, a record for one read in three, each with up to 5 edits or a full
    sequence, handed over in reverse order."""
    n = 200_000
    data = fastq_data.make(seed=21, n_reads=n, paired=False, min_len=100, max_len=150)
    rng = np.random.default_rng(4)
    recs = []
    for i in range(0, n, 3):
        L = int(data["lengths"][i])
        if rng.random() < 0.8:
            k = int(rng.integers(0, 6))
            pos = np.sort(rng.choice(L, k, replace=False))
            edits = [(int(p), b"ACGT"[int(rng.integers(0, 4))].to_bytes(1, "little")) for p in pos]
            recs.append(encode_record(i, hq=bool(rng.random() < 0.5), edits=edits))
        else:
            seq = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)])
            recs.append(encode_record(i, hq=True, use_edits=False, sequence=seq))
    records = b"".join(reversed(recs))
    text = fastq_data.text(data, range(n))
    for labels in (False, True):
        got = _gpu_merge([text], records, False, labels)
        assert got == output_merge.merge([text], records, False, labels)


def test_merge_rejects_bad_input():
    data = fastq_data.make(seed=2, n_reads=50, paired=False)
    text = fastq_data.text(data, range(50))
    eng = engine.CareEngine(0, paired=0, use_quality_scores=1)
    try:
        eng.set_reads_fastq(text)
        with pytest.raises(engine.CareError, match="read id"):
            eng.write_fastq(encode_record(50, edits=[(0, b"A")]))
        with pytest.raises(engine.CareError, match="truncated"):
            eng.write_fastq(encode_record(3, edits=[(0, b"A")])[:-1])
        bad = struct.pack("<II", 3, (1 << 30) | 4) + struct.pack("<i", 7)      # 7 edits, no room
        with pytest.raises(engine.CareError, match="malformed"):
            eng.write_fastq(bad)
        with pytest.raises(engine.CareError, match="out of range"):
            eng.write_fastq(b"", file=1)
    finally:
        eng.close()
    s, rs, o = helpers.case("se")
    eng = engine.CareEngine(0, **o)
    try:
        eng.set_reads(rs)
        with pytest.raises(engine.CareError, match="set_reads_fastq"):
            eng.write_fastq(b"")
    finally:
        eng.close()
