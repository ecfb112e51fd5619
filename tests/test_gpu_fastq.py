"""Read storage built on the GPU from FASTQ text (care_hip_set_reads_fastq), and the
preprocessed-reads file (care_hip_save_reads / care_hip_load_reads).

Bar: bit-exact. The device storage (2-bit rows, lengths, ambiguity flags, qualities) equals
the host construction rabbitqcplus_amd.reads.make_storage, which the CPU suite pins to the
reference's own FASTQ -> storage construction on the same seeded data (more than one
parser batch of 65,536 reads, N and IUPAC codes, lower case); corrections from the
FASTQ-built storage equal the golden records of the same reads. The engine's reads file
equals the reference's file (the ambiguous-id list compared in sorted order, which the
reference writes in its hash set's order), and the files load both ways.
"""
import os

import numpy as np
import pytest

import oracle
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


def _ambig_sorted(path):
    """File bytes with the ambiguous-id list sorted (the reference writes its hash set's order)."""
    f = fastq_data.parse_reads_file(path)
    blob = open(path, "rb").read()
    na = len(f["ambiguous_ids"])
    return blob[:len(blob) - 4 * na] + np.sort(f["ambiguous_ids"]).astype("<u4").tobytes()


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("paired, use_q", [(False, True), (True, True), (False, False)])
def test_reads_file_equals_reference_and_loads_both_ways(tmp_path, paired, use_q):
    data = fastq_data.make(seed=12 + paired, n_reads=70_000 if not paired else 2 * 36_000, paired=paired)
    files = fastq_data.write(tmp_path, data, paired)
    theirs = str(tmp_path / "ref.bin")
    ref = oracle.ref_parse_fastq(*files, use_qualities=use_q, save_to=theirs)
    print(f"[reads file paired={paired} q={use_q}] reads={len(ref['lengths'])}", flush=True)
    texts = _texts(data, paired)
    ours = str(tmp_path / "ours.bin")
    eng = engine.CareEngine(0, paired=int(paired), use_quality_scores=int(use_q))
    try:
        eng.set_reads_fastq(*texts)
        eng.save_reads(ours)
    finally:
        eng.close()
    assert np.array_equal(np.sort(fastq_data.parse_reads_file(ours)["ambiguous_ids"]),
                          fastq_data.parse_reads_file(ours)["ambiguous_ids"])
    assert _ambig_sorted(ours) == _ambig_sorted(theirs)
    # the reference's file into the engine
    rs = _host_storage(data, paired, use_q)
    eng = engine.CareEngine(0, paired=int(paired), use_quality_scores=int(use_q))
    try:
        assert eng.load_reads(theirs) == rs.n
        _assert_same(eng.inspect_reads(), rs, use_q)
    finally:
        eng.close()
    # the engine's file into the reference
    back = oracle.ref_load_reads(ours, paired, use_q)
    assert np.array_equal(back["enc"], rs.enc) and np.array_equal(back["lengths"], rs.lengths)
    assert np.array_equal(back["ambiguous"], rs.ambiguous)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_records_from_loaded_reads_equal_golden(tmp_path):
    s, rs, o = helpers.case("se_q")
    data = dict(seq_matrix=s.seqs, qual_matrix=s.quals, lengths=s.lengths)
    files = fastq_data.write(tmp_path, data, False)
    path = str(tmp_path / "ref.bin")
    oracle.ref_parse_fastq(*files, use_qualities=True, save_to=path)
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", "records_se_q.npz"))["records"].tobytes()
    eng = engine.CareEngine(0, **o)
    try:
        eng.load_reads(path)
        eng.build_tables()
        assert eng.correct()[1] == golden
    finally:
        eng.close()


def test_reads_file_errors(tmp_path):
    data = fastq_data.make(seed=30, n_reads=500, paired=False)
    text = _texts(data, False)[0]
    noq = str(tmp_path / "noq.bin")
    eng = engine.CareEngine(0, paired=0, use_quality_scores=0)
    try:
        eng.set_reads_fastq(text)
        eng.save_reads(noq)
    finally:
        eng.close()
    eng = engine.CareEngine(0, paired=0, use_quality_scores=1)
    try:
        with pytest.raises(engine.CareError, match="quality scores expected"):
            eng.load_reads(noq)
        eng.set_reads_fastq(text)
        withq = str(tmp_path / "q.bin")
        eng.save_reads(withq)
        for cut in (1000, 30000):
            short = str(tmp_path / "short.bin")
            open(short, "wb").write(open(withq, "rb").read()[:cut])
            with pytest.raises(engine.CareError, match="truncated|larger than the file"):
                eng.load_reads(short)
    finally:
        eng.close()


def _reads_file_header(raw: bytes):
    """Offsets of the length section of a preprocessed-reads file: bits per length at 57,
    minimum and maximum length at 61 and 65, word count at 81, the words from 97."""
    bits, min_len, max_len = (int.from_bytes(raw[o:o + 4], "little", signed=True) for o in (57, 61, 65))
    lwords = int.from_bytes(raw[81:89], "little")
    return bits, min_len, max_len, lwords


def test_reads_file_with_bad_sizes_is_rejected_before_any_allocation(tmp_path):
    """A length stream whose values exceed the header's maximum length, and sections whose
    sizes exceed the file, are CARE_EINVAL with a message (the longest read sizes the rows,
    the MSA bounds and the record slots; sizes are checked before allocating)."""
    data = fastq_data.make(seed=31, n_reads=800, paired=False)
    eng = engine.CareEngine(0, paired=0, use_quality_scores=1)
    try:
        eng.set_reads_fastq(_texts(data, False)[0])
        good = str(tmp_path / "good.bin")
        eng.save_reads(good)
        raw = open(good, "rb").read()
        bits, min_len, max_len, lwords = _reads_file_header(raw)
        assert bits > 0 and min_len + (1 << bits) - 1 > max_len, "the workload needs variable lengths"
        cases = {
            # every length value all ones: min + 2^bits - 1 > max
            "longer": raw[:97] + b"\xff" * (4 * lwords) + raw[97 + 4 * lwords:],
            # word count (and byte count) far beyond the file
            "larger than the file": raw[:81] + (1 << 40).to_bytes(8, "little") + (4 << 40).to_bytes(8, "little")
                                    + raw[97:],
        }
        # sequence section: pitch kept, row count claimed 2^30 times larger
        sp_at = 97 + 4 * lwords
        sp = int.from_bytes(raw[sp_at:sp_at + 8], "little")
        n = int.from_bytes(raw[0:8], "little")
        cases["bad sequence section|larger than the file"] = (raw[:sp_at + 8] + (sp * n * (1 << 30)).to_bytes(8, "little")
                                                             + raw[sp_at + 16:])
        for i, (match, blob) in enumerate(cases.items()):
            path = str(tmp_path / f"bad{i}.bin")
            open(path, "wb").write(blob)
            with pytest.raises(engine.CareError, match=match):
                eng.load_reads(path)
        # the context still works after the rejections
        assert eng.load_reads(good) == n
    finally:
        eng.close()
