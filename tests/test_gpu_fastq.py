"""Read storage built on the GPU from FASTQ text (care_hip_set_reads_fastq).

Bar: bit-exact. The device storage (2-bit rows, lengths, ambiguity flags, qualities) equals
the host construction rabbitqcplus_amd.reads.make_storage, which the CPU suite pins to the
reference's own FASTQ -> storage construction on the same seeded data (more than one
parser batch of 65,536 reads, N and IUPAC codes, lower case); corrections from the
FASTQ-built storage equal the golden records of the same reads.
"""
import os

import numpy as np
import pytest

from rabbitqcplus_amd import engine, reads
from tests import fastq_data, helpers

pytestmark = pytest.mark.gpu


def _gpu_storage(texts, paired, use_q=True):
    eng = engine.CareEngine(0, paired=int(paired), use_quality_scores=int(use_q))
    try:
        n = eng.set_reads_fastq(*texts)
        got = eng.inspect_reads()
        assert got["enc"].shape[0] == n
    finally:
        eng.close()
    return got


def _host_storage(data, paired, use_q=True):
    return reads.make_storage(data["seq_matrix"], data["lengths"], data["qual_matrix"] if use_q else None, paired)


def _texts(data, paired, mangle=None):
    n = len(data["lengths"])
    parts = [range(n)] if not paired else [range(0, n, 2), range(1, n, 2)]
    out = [fastq_data.text(data, p) for p in parts]
    return [mangle(t) for t in out] if mangle else out


def _assert_same(got, rs, use_q=True):
    assert np.array_equal(got["lengths"], rs.lengths)
    assert np.array_equal(got["ambiguous"], rs.ambiguous)
    assert got["enc"].shape == rs.enc.shape
    assert np.array_equal(got["enc"], rs.enc)
    if use_q:
        ml = got["max_length"]
        valid = np.arange(ml)[None, :] < rs.lengths[:, None]
        assert np.array_equal(np.where(valid, got["quals"], 0), np.where(valid, rs.quals[:, :ml], 0))
        assert not np.where(valid, 0, got["quals"]).any()     # zero past each read's end
    else:
        assert got["quals"] is None


@pytest.mark.parametrize("paired", [False, True])
def test_fastq_storage_matches_host(paired):
    data = fastq_data.make(seed=5 + paired, n_reads=70_000 if not paired else 2 * 36_000, paired=paired)
    print(f"[fastq paired={paired}] reads={len(data['lengths'])}", flush=True)
    rs = _host_storage(data, paired)
    assert rs.ambiguous.sum() > 100
    _assert_same(_gpu_storage(_texts(data, paired), paired), rs)


def test_fastq_crlf_separator_text_and_trailing_blank_lines():
    data = fastq_data.make(seed=9, n_reads=3000, paired=False, min_len=1, max_len=300)

    def mangle(t):
        t = t.replace(b"\n+\n", b"\n+ copy of the header\n").replace(b"\n", b"\r\n")
        return t + b"\r\n\n\n"

    rs = _host_storage(data, False)
    _assert_same(_gpu_storage(_texts(data, False, mangle), False), rs)


def test_fastq_without_qualities_and_without_final_newline():
    data = fastq_data.make(seed=10, n_reads=5000, paired=False)
    t = _texts(data, False)[0].rstrip(b"\n")
    _assert_same(_gpu_storage([t], False, use_q=False), _host_storage(data, False, use_q=False), use_q=False)


@pytest.mark.parametrize("name", ["se_q", "pe", "varlen_q"])
def test_records_from_fastq_equal_golden(name):
    s, rs, o = helpers.case(name)
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", f"records_{name}.npz"))["records"].tobytes()
    data = dict(seq_matrix=s.seqs, qual_matrix=s.quals, lengths=s.lengths)
    texts = _texts(data, rs.paired)
    print(f"[case {name}] reads={rs.n}", flush=True)
    eng = engine.CareEngine(0, **o)
    try:
        assert eng.set_reads_fastq(*texts) == rs.n
        _assert_same(eng.inspect_reads(), rs, use_q=bool(o.get("use_quality_scores")))
        eng.build_tables()
        assert eng.correct()[1] == golden
    finally:
        eng.close()


@pytest.mark.parametrize("bad, match", [
    (b"@r0\nACGT\n+\nIIII\n@r1\nACGT\n", "4-line"),
    (b"@r0\nACGT\n-\nIIII\n", "malformed"),
    (b"r0\nACGT\n+\nIIII\n", "malformed"),
    (b"@r0\nACGT\n+\nIII\n", "malformed"),
    (b">r0\nACGT\n>r1\nACGT\n", "malformed"),
    (b"@r0\nACGT\nACGT\n+\nIIII\nIIII\n@r1\nACGT\nACGT\n+\nIIII\nIIII\n", "malformed"),
    (b"\n\n", "no FASTQ records|4-line|empty"),
])
def test_fastq_rejects_malformed_text(bad, match):
    eng = engine.CareEngine(0, paired=0, use_quality_scores=1)
    try:
        with pytest.raises(engine.CareError, match=match):
            eng.set_reads_fastq(bad)
    finally:
        eng.close()


def test_paired_fastq_needs_two_texts_of_equal_records():
    rec = b"@r\nACGT\n+\nIIII\n"
    eng = engine.CareEngine(0, paired=1, use_quality_scores=1)
    try:
        with pytest.raises(engine.CareError, match="two FASTQ texts"):
            eng.set_reads_fastq(rec * 2)
        with pytest.raises(engine.CareError, match="records"):
            eng.set_reads_fastq(rec * 2, rec * 3)
        assert eng.set_reads_fastq(rec * 2, rec * 2) == 4
    finally:
        eng.close()
