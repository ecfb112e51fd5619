"""Records at the benchmark's full sizes against the reference engine compiled from source.

For each workload of helpers.FULL_WORKLOADS (1.0M SE reads of 150 bp; 2.0M pairs of
2x150 bp with 20-copy repeat families; 8.57M SE reads of 100..250 bp with 10-copy
repeats, k = 20), tests/golden/full_<workload>.json holds the reference's records
(qualities on) as record counts and SHA-256 digests over 16 read-id ranges, written by
tests/golden/make_golden.py in the build container, with the SHA-256 of the simulated
storage they belong to.

Bar: bit-exact. Properties that need no reference: the records come in ascending read-id
order, and do not depend on the batch size or on how the anchor range is split. The fast
learned form (learned seam on seeded alignment) leaves no more errors than the plain path.
"""
import json
import os

import numpy as np
import pytest

from rabbitqcplus_amd import engine, evaluate, learned
from tests import helpers
from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WORKLOADS = list(helpers.FULL_WORKLOADS)


def _golden(w):
    with open(os.path.join(HERE, "golden", f"full_{w}.json")) as fh:
        return json.load(fh)


_cache = {}


def _case(w):
    if w not in _cache:
        _cache.clear()                      # one full-size workload in memory at a time
        s, rs, o = helpers.full_case(w)
        _cache[w] = (s, rs, o)
    return _cache[w]


def _pass(rs, o, **extra):
    eng = engine.CareEngine(0, **o, **extra)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        return eng.correct()
    finally:
        eng.close()


@pytest.mark.parametrize("w", WORKLOADS)
def test_full_size_records_equal_reference(w):
    g = _golden(w)
    s, rs, o = _case(w)
    print(f"[full {w}] reads={rs.n}", flush=True)
    assert rs.n == g["reads"] and mg.storage_digest(rs) == g["storage_sha256"], "simulated input differs"
    n, data = _pass(rs, o)
    assert n == g["records"] and len(data) == g["record_bytes"]
    got = helpers.slice_digests(data, rs.n)
    bad = [i for i, (x, y) in enumerate(zip(got, g["slices"])) if x != y]
    assert not bad, f"read-id ranges differing from the reference: {bad}"


@pytest.mark.parametrize("w", WORKLOADS)
def test_full_size_batch_and_range_invariance(w):
    s, rs, o = _case(w)
    print(f"[full {w}] invariance", flush=True)
    want = _pass(rs, o)[1]
    B = (rs.n // 3) & ~1
    assert _pass(rs, o, batch_size=B)[1] == want
    eng = engine.CareEngine(0, **o)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        k = (rs.n * 2 // 5) & ~1
        assert eng.correct(0, k)[1] + eng.correct(k, rs.n - k)[1] == want
    finally:
        eng.close()


def _record_count(data):
    mv = memoryview(data)
    p = n = 0
    while p < len(data):
        p += 8 + (int.from_bytes(mv[p + 4:p + 8], "little") & ((1 << 29) - 1))
        n += 1
    assert p == len(data)
    return n


def test_full_size_candidate_correction_batch_invariance():
    """Candidate correction over the 8.57M-read workload at the automatic batch size (about
    2.7M anchors and 165M working-list entries per batch: the record gather strides over
    more than 2^32 work items' worth of entries) gives the same record stream as batches of
    2^20 anchors, and the reported record count is the number of records in it."""
    w = "se_50mbp_100to250bp_30x_rep"
    s, rs, o = _case(w)
    outs = []
    for bs in (0, 1 << 20):
        eng = engine.CareEngine(0, correct_candidates=1, batch_size=bs, **o)
        try:
            eng.set_reads(rs)
            eng.build_tables()
            n, data = eng.correct()
            outs.append((n, data, eng.timings()["batches"]))
        finally:
            eng.close()
        print(f"[full {w}] candidate correction, batch size {bs}: {outs[-1][0]} records in {outs[-1][2]} batches",
              flush=True)
    assert outs[0][2] < outs[1][2]
    assert outs[0][1] == outs[1][1]
    assert outs[0][0] == outs[1][0] == _record_count(outs[0][1])


def _record_index(data):
    """(read ids, byte offsets) of the records of a stream, in stream order."""
    import struct
    ids, offs = [], []
    p, end = 0, len(data)
    while p < end:
        rid, flags = struct.unpack_from("<II", data, p)
        ids.append(rid)
        offs.append(p)
        p += 8 + (flags & ((1 << 29) - 1))
    assert p == end
    return np.asarray(ids, np.int64), np.asarray(offs, np.int64)


def _lines(buf, nl, i):
    """The 4 lines of read i of 4-line FASTQ text (newline positions nl)."""
    starts = [0 if 4 * i == 0 else int(nl[4 * i - 1]) + 1] + [int(nl[4 * i + j]) + 1 for j in range(3)]
    return [bytes(buf[a:int(nl[4 * i + j])]) for j, a in enumerate(starts)]


def test_full_size_fastq_text_to_corrected_text():
    """FASTQ text of the 8.57M-read workload (about 3 GB) -> device read storage ->
    correction -> corrected FASTQ text, with and without quality labels. The merge writes
    with one wavefront per read, so its grid-stride loop takes several strides (8.57M reads
    x 64 lanes > 2^28 work items of the largest grid).

    Checks: the records from the text equal the reference's (the golden digests); without
    labels the output is the input with bytes changed only inside sequence lines of reads
    that have a record; 20,000 random reads and the reads either side of each grid stride
    equal the oracle's combine rule applied to their records, with and without labels."""
    from oracle import output_merge
    from rabbitqcplus_amd import sim
    w = "se_50mbp_100to250bp_30x_rep"
    s, rs, o = _case(w)
    text = sim.fastq_texts(s)[0]
    print(f"[full {w}] FASTQ text {text.size / 1e9:.2f} GB", flush=True)
    eng = engine.CareEngine(0, **o)
    try:
        assert eng.set_reads_fastq(text) == rs.n
        eng.build_tables()
        n, data = eng.correct()
        # (bytes, and the array form the bench uses)
        outs = [np.frombuffer(eng.write_fastq(data, 0, False), np.uint8), eng.write_fastq(data, 0, True, as_array=True)]
    finally:
        eng.close()
    g = _golden(w)
    assert n == g["records"] and helpers.slice_digests(data, rs.n) == g["slices"]
    ids, offs = _record_index(data)
    assert np.all(np.diff(ids) >= 0)
    nl = np.flatnonzero(text == 10)
    assert nl.size == 4 * rs.n
    plain, labelled = outs
    # equal-length corrections: the unlabelled output lines up with the input byte for byte
    assert plain.size == text.size
    diff = np.flatnonzero(plain != text)
    line = np.searchsorted(nl, diff)
    assert np.all(line % 4 == 1) and np.all(text[diff] != 10) and np.all(plain[diff] != 10)
    assert np.isin(np.unique(line // 4), ids).all()
    print(f"[full {w}] {diff.size} bytes corrected in {np.unique(line // 4).size} reads", flush=True)
    nll = np.flatnonzero(labelled == 10)
    assert nll.size == 4 * rs.n
    per_stride = (1 << 28) // 64                       # reads per stride of the merge's grid
    edge = [k * per_stride + d for k in range(1, rs.n // per_stride + 1) for d in (-1, 0)]
    sample = np.unique(np.concatenate([np.random.default_rng(5).choice(rs.n, 20000, replace=False),
                                       [0, rs.n - 1], [e for e in edge if e < rs.n]]))
    last = int(ids[-1])
    tag = {0: b" care:q=0", 1: b" care:q=1", 2: b" care:q=2", 3: b" care:q=3"}
    for i in sample:
        i = int(i)
        h, seq, plus, q = _lines(text, nl, i)
        a, b = np.searchsorted(ids, i), np.searchsorted(ids, i, side="right")
        recs = [output_merge.decode_records(data[int(offs[k]):int(offs[k + 1]) if k + 1 < len(offs) else len(data)])[0]
                for k in range(a, b)]
        want, lab = output_merge.combine(recs, seq)
        assert _lines(plain, nl, i) == [h, want, plus, q], i
        hl = h + tag[lab] if i <= last else h
        assert _lines(labelled, nll, i) == [hl, want, plus, q], i


@pytest.mark.parametrize("w", ["pe_20mbp_2x150bp_30x_rep", "se_50mbp_100to250bp_30x_rep"])
def test_full_size_fast_learned_form_not_worse_than_plain(w):
    s, rs, o = _case(w)
    params, _ = learned.load_classifier()
    _, plain = _pass(rs, o)
    eng = engine.CareEngine(0, correction_type=1, seeded_alignment=1, **o)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        eng.set_classifier(params)
        fast = eng.correct()[1]
    finally:
        eng.close()
    ep, ef = evaluate.error_counts(s, rs, plain), evaluate.error_counts(s, rs, fast)
    print(f"[full {w}] plain {ep} / fast learned {ef}", flush=True)
    assert ef["after"] <= ep["after"]


def test_training_on_200kbp_columns_at_batch_65536():
    """The training configuration: decision columns (16 features) of the 200 kbp /
    150 bp / 30x workload, labelled by the simulator's truth, MLP 16-64-64-1 trained on the
    GPU with Adam (lr 1e-3) at batch 65,536. The device gradient of a full 65,536-row batch
    agrees with the oracle's double-precision backpropagation, and training lowers the loss."""
    import oracle
    from rabbitqcplus_amd import reads, sim
    # the 200 kbp workload yields about 115 decision columns per 40k reads: the same read
    # profile (uniform genome, 150 bp, 30x, Q36 -> Q28) over a 120 Mbp genome (24M reads)
    # gives a full batch
    kw = dict(sim.WORKLOADS["se_200kbp_150bp_30x"], genome_length=120_000_000, per_read_streams=1, seed=1000)
    s = sim.simulate(**kw)
    rs = reads.storage_from_sim(s, use_qualities=True)
    X, y, _ = learned.collect(s, rs, use_quality_scores=1, paired=0, kmerlength=20)
    del s, rs
    assert len(y) >= 65536, len(y)
    X, y = X[:65536], y[:65536]
    print(f"{len(y)} columns; positives {y.mean():.4f}", flush=True)
    mean, std = X.mean(axis=0), X.std(axis=0) + 1e-6
    Xn = ((X - mean) / std).astype(np.float32)
    p = learned.init_params(0)
    mlp = learned.Mlp(0)
    try:
        mlp.set_params(p)
        loss, g = mlp.loss_grad(Xn, y)
        loss64, g64, _ = oracle.mlp_loss_grad_f64(p.astype(np.float64), Xn, y)
        scale = np.abs(g64).max()
        assert abs(loss - loss64) <= 1e-5 * max(1.0, abs(loss64))
        assert np.abs(g - g64).max() <= 1e-4 * scale
        mlp.set_data(Xn, y)
        losses = mlp.train(epochs=40, batch=65536, lr=1e-3, seed=0)
    finally:
        mlp.close()
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.8 * losses[0]
