"""Pipelined staged batches (no host synchronization inside a batch), with and without
HIP graphs, against the per-batch-synchronized schedule and the oracle.

The pipelined schedule sizes a batch's candidate arrays before its candidates are counted;
a batch with more candidates is redone with exact sizes. Bar: bit-exact records, whether
batches are redone or not, whatever the batch size, and whether each batch is a captured
graph replayed across passes.
"""
import ctypes

import pytest

import oracle
from rabbitqcplus_amd import engine, learned
from tests import helpers

pytestmark = pytest.mark.gpu


def _engine(rs, o, **extra):
    eng = engine.CareEngine(0, **{**o, **extra})
    eng.set_reads(rs)
    eng.build_tables()
    return eng


def _oracle_records(rs, o, count=False):
    cpu = oracle.CareCpu(rs, threads=8, **o)
    try:
        n, data = cpu.correct(threads=8)
        return (n, data) if count else data
    finally:
        cpu.close()


@pytest.mark.parametrize("name", ["se_q", "pe", "varlen_q", "se_higherr", "len300_q", "nbases_exclude"])
def test_pipelined_matches_synchronized_and_oracle(name):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    nwant, want = _oracle_records(rs, o, count=True)
    for extra in ({}, {"batch_size": 1000}, {"batch_size": 1000, "use_graphs": 1}, {"sync_batches": 1}):
        eng = _engine(rs, o, engine=1, **extra)
        try:
            n, d = eng.correct()
            t = eng.timings()
        finally:
            eng.close()
        assert d == want, extra
        assert n == nwant == t["records"], extra
        if "batch_size" in extra:
            assert t["batches"] == -(-rs.n // 1000)
        if extra.get("use_graphs"):
            assert t["graph_ms"] > 0 and t["graphs_built"] == t["batches"]


@pytest.mark.parametrize("graphs", [0, 1])
def test_batches_over_capacity_are_redone(graphs):
    """A first capacity of one pair per anchor (the case has about 46): the first two
    batches (both enqueued before the first is counted) are redone with exact sizes, the
    later ones are sized from the counts seen and run pipelined; the records equal the
    oracle's. A second pass on the context redoes none."""
    s, rs, o = helpers.case("se_higherr_q")
    want = _oracle_records(rs, o)
    eng = _engine(rs, o, engine=1, batch_size=777, pipeline_capacity=1, use_graphs=graphs)
    try:
        d = eng.correct()[1]
        t = eng.timings()
        print(f"graphs {graphs}: {t['batches_redone']} of {t['batches']} batches redone", flush=True)
        assert d == want
        assert t["batches"] == -(-rs.n // 777) and t["batches_redone"] == 2
        assert eng.correct()[1] == want
        assert eng.timings()["batches_redone"] == 0
    finally:
        eng.close()


def test_staged_graph_replays_match_eager():
    """Each pipelined batch captured once and replayed on later passes: every replay
    equals the eager pass, with eager work on another context between replays; a new
    anchor range captures new graphs."""
    s, rs, o = helpers.case("se_higherr_q")
    eager = _engine(rs, o, engine=1, batch_size=2048)
    want = eager.correct()[1]
    g = _engine(rs, o, engine=1, batch_size=2048, use_graphs=1)
    try:
        assert g.correct()[1] == want
        built = g.timings()["graphs_built"]
        assert built == -(-rs.n // 2048)
        for rep in range(3):
            assert eager.correct(0, 3000)[1] == eager.correct(0, 3000)[1]
            assert g.correct()[1] == want, f"replay {rep}"
            assert g.timings()["graphs_built"] == built
        m = 1234
        assert g.correct(0, m)[1] + g.correct(m, rs.n - m)[1] == want
        assert g.timings()["graphs_built"] > built
    finally:
        g.close()
        eager.close()


def test_seeded_alignment_pipelined_matches_synchronized():
    s, rs, o = helpers.case("varlen_q")
    cpu = oracle.CareCpu(rs, threads=8, **o)
    try:
        cpu.set_seeded_alignment(True)
        want = cpu.correct(threads=8)[1]
    finally:
        cpu.close()
    for extra in ({}, {"use_graphs": 1, "batch_size": 1500}, {"sync_batches": 1}):
        eng = _engine(rs, o, engine=1, seeded_alignment=1, **extra)
        try:
            assert eng.correct()[1] == want, extra
        finally:
            eng.close()


def test_learned_pipelined_matches_synchronized():
    params, _ = learned.load_classifier()
    s, rs, o = helpers.case("se_higherr_q")
    outs = []
    for extra in ({}, {"use_graphs": 1, "batch_size": 1500}, {"sync_batches": 1}):
        eng = _engine(rs, o, engine=1, correction_type=1, **extra)
        try:
            eng.set_classifier(params)
            outs.append(eng.correct()[1])
        finally:
            eng.close()
    assert outs[0] == outs[1] == outs[2]
    assert len(outs[0]) > 0


def test_records_view_equals_copy():
    """care_hip_records_view (what CareEngine.correct reads) and care_hip_copy_records
    give the same bytes; the record count matches the records in them."""
    s, rs, o = helpers.case("se_q")
    eng = _engine(rs, o)
    try:
        n, data = eng.correct()
        assert n == len(helpers.parse_records(data)) > 0
        buf = ctypes.create_string_buffer(len(data))
        assert eng.lib.care_hip_copy_records(eng.h, buf, len(data)) == 0
        assert buf.raw == data
        assert eng.lib.care_hip_copy_records(eng.h, buf, len(data) - 1) != 0
    finally:
        eng.close()


@pytest.mark.parametrize("extra", [{}, {"use_graphs": 1}, {"sync_batches": 1}, {"engine": 2}])
def test_empty_and_small_anchor_ranges(extra):
    """Ranges of no anchor, one anchor and a few anchors: the records of exactly those
    anchors, as the whole pass gives them (records come in read-id order)."""
    s, rs, o = helpers.case("se_q")
    eng = _engine(rs, o, **extra)
    try:
        whole = helpers.parse_records(eng.correct()[1])
        assert eng.correct(0, 0) == (0, b"")
        assert eng.correct(rs.n, 0) == (0, b"")
        for first, n in ((rs.n - 1, 1), (0, 1), (17, 5), (1000, 63), (2000, 65)):
            nrec, data = eng.correct(first, n)
            got = helpers.parse_records(data)
            want = [r for r in whole if first <= r[0] < first + n]
            assert got == want and nrec == len(want), (first, n)
    finally:
        eng.close()
